"""Seeded sweep over random Leopard shapes, per call and batched.

Shapes are drawn over the whole valid range (GF(2^8): k + ceilPow2(m) <= 256;
GF(2^16): up to 4096 total shards at small sizes), so the encode groups, the
truncations, the slice widths and the schedules the kernels pick vary from
stripe to stripe.  Parity must equal the oracle's (a restatement of
leopard.go / leopard8.go encode); a reconstruct of a valid codeword has one
answer, the original shards.
"""
import numpy as np
import pytest

from oracle import rsref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import reedsolomon_amd as rs  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available() or rs.device_count() == 0:
        pytest.skip("no GPU")


def ceil_pow2(x):
    return 1 << (x - 1).bit_length()


def draw_shapes(rng, bits, count):
    shapes = []
    while len(shapes) < count:
        if bits == 8:
            m = int(rng.integers(1, 129))
            kmax = 256 - ceil_pow2(m)
            if kmax < 1:
                continue
            k = int(rng.integers(1, kmax + 1))
        else:
            m = int(rng.integers(1, 1025))
            k = int(rng.integers(1, 4097 - m))
        shapes.append((k, m, 64 * int(rng.integers(1, 7 if bits == 8 else 4))))
    return shapes


def codeword(rng, bits, k, m, size):
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    shards = data + [np.zeros(size, np.uint8) for _ in range(m)]
    assert rsref.Leopard(k, m, bits).encode(shards) == 0
    return shards


def opt(bits):
    return rs.WithLeopardGF16(True) if bits == 16 else rs.WithLeopardGF(True)


@pytest.mark.parametrize("bits,count,seed", [(8, 40, 81), (16, 16, 161)])
def test_random_shapes_per_call_and_batched(bits, count, seed):
    rng = np.random.default_rng(seed)
    shapes = draw_shapes(rng, bits, count)
    items, checks = [], []
    for n, (k, m, size) in enumerate(shapes):
        case = (bits, k, m, size)
        enc = rs.New(k, m, opt(bits))
        full = codeword(rng, bits, k, m, size)
        dev = [torch.from_numpy(x).to("cuda") for x in full]
        # Per call: encode into zeroed parity, then rebuild 1..m erased shards.
        shards = [x.clone() for x in dev[:k]] + [torch.zeros(size, dtype=torch.uint8, device="cuda")
                                                 for _ in range(m)]
        enc.Encode(shards)
        for i in range(k, k + m):
            assert np.array_equal(shards[i].cpu().numpy(), full[i]), (case, "encode", i)
        lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
        work = [None if i in lost else dev[i].clone() for i in range(k + m)]
        enc.Reconstruct(work)
        for i in sorted(lost):
            assert np.array_equal(work[i].cpu().numpy(), full[i]), (case, "reconstruct", i, len(lost))
        # The same stripe as a batch item: even n encode into stale parity,
        # odd n reconstruct another erasure pattern.
        if n % 2 == 0:
            stale = [x.clone() for x in dev[:k]] + [
                torch.from_numpy(rng.integers(0, 256, size, dtype=np.uint8)).to("cuda") for _ in range(m)]
            items.append((enc, rs.OP_ENCODE, stale))
            checks.append((case, full, stale, range(k, k + m)))
        else:
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            work = [None if i in lost else dev[i].clone() for i in range(k + m)]
            items.append((enc, rs.OP_RECONSTRUCT, work))
            checks.append((case, full, work, sorted(lost)))
    # Aligned device shards of valid shapes: every item runs in the plan.
    b = rs.Batch(0)
    status = b.run(items)
    assert status == [0] * len(items), status
    for case, full, shards, rows in checks:
        for i in rows:
            assert np.array_equal(shards[i].cpu().numpy(), full[i]), (case, "batched", i)
