"""Seeded sweep over random classic GF(2^8) stripes, per call and batched.

Each stripe draws its shape over the whole range (k + m <= 256, so m > 8
takes several output passes and the batch splits it into 8-output tasks),
an odd shard size, a byte offset (unaligned shards take the byte paths) and
the default or Cauchy matrix.  Parity must equal the oracle's encode
(reedsolomon.go:686 restated); a reconstruct of a valid codeword has one
answer, the original shards (both matrices are MDS).
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


def stripe_views(k, m, size, offset):
    """k + m device shards of `size` bytes starting `offset` bytes into
    consecutive (size + offset)-byte runs of one allocation."""
    pool = torch.zeros((k + m) * (size + offset) + 16, dtype=torch.uint8, device="cuda")
    return [pool[offset + i * (size + offset): offset + i * (size + offset) + size] for i in range(k + m)]


@pytest.mark.parametrize("seed", [7, 8])
def test_random_stripes_per_call_and_batched(seed):
    rng = np.random.default_rng(seed)
    items, checks = [], []
    for n in range(32):
        m = int(rng.integers(1, 41))
        k = int(rng.integers(1, 257 - m))
        size = int(rng.integers(1, 6000))
        offset = int(rng.integers(0, 16)) if n % 2 else 0
        cauchy = bool(rng.integers(0, 2))
        case = (k, m, size, offset, "cauchy" if cauchy else "default")
        kind = rsref.MATRIX_CAUCHY if cauchy else rsref.MATRIX_DEFAULT
        full = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + [
            np.zeros(size, np.uint8) for _ in range(m)]
        assert rsref.Classic(k, m, kind).encode(full) == 0
        enc = rs.New(k, m, rs.WithCauchyMatrix()) if cauchy else rs.New(k, m)
        # Per call: encode into stale parity, then rebuild 1..m erased shards.
        shards = stripe_views(k, m, size, offset)
        for i in range(k):
            shards[i].copy_(torch.from_numpy(full[i]))
        for i in range(k, k + m):
            shards[i].copy_(torch.from_numpy(rng.integers(0, 256, size, dtype=np.uint8)))
        enc.Encode(shards)
        for i in range(k, k + m):
            assert np.array_equal(shards[i].cpu().numpy(), full[i]), (case, "encode", i)
        lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
        work = [None if i in lost else shards[i] for i in range(k + m)]
        enc.Reconstruct(work)
        for i in sorted(lost):
            assert np.array_equal(work[i].cpu().numpy(), full[i]), (case, "reconstruct", i, len(lost))
        # The same stripe as a batch item in fresh buffers.
        views = stripe_views(k, m, size, offset)
        for i in range(k + m):
            views[i].copy_(torch.from_numpy(full[i] if i < k else rng.integers(0, 256, size, dtype=np.uint8)))
        if n % 3 != 2:
            items.append((enc, rs.OP_ENCODE, views))
            checks.append((case, full, views, range(k, k + m)))
        else:
            for i in range(k, k + m):
                views[i].copy_(torch.from_numpy(full[i]))
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            work = [None if i in lost else views[i] for i in range(k + m)]
            items.append((enc, rs.OP_RECONSTRUCT, work))
            checks.append((case, full, work, sorted(lost)))
    b = rs.Batch(0)
    status = b.run(items)
    assert status == [0] * len(items), status
    for case, full, shards, rows in checks:
        for i in rows:
            assert np.array_equal(shards[i].cpu().numpy(), full[i]), (case, "batched", i)
