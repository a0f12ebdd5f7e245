"""GPU parity of the batched path: every item of a mixed plan must equal the
plain per-stripe call (and the oracle) byte for byte, and report the same
error the plain call reports."""
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


SHAPES = [(4, 2), (6, 3), (10, 4), (17, 3), (32, 8), (20, 12)]


def _stripe(rng, k, m, size, offset=0):
    pool = torch.zeros((k + m) * (size + offset) + 16, dtype=torch.uint8, device="cuda")
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    shards = []
    for i in range(k + m):
        t = pool[offset + i * (size + offset): offset + i * (size + offset) + size]
        if i < k:
            t.copy_(torch.from_numpy(data[i]))
        shards.append(t)
    return data, shards


def test_mixed_batch_matches_plain_and_oracle():
    rng = np.random.default_rng(2024)
    codecs = {s: rs.New(*s) for s in SHAPES}
    items, expect = [], []
    for n in range(60):
        k, m = SHAPES[n % len(SHAPES)]
        size = int(rng.choice([16, 64, 1000, 4096, 5000, 70000, 1 << 18]))
        offset = 1 if n % 7 == 0 else 0  # some unaligned stripes
        data, shards = _stripe(rng, k, m, size, offset)
        o = rsref.Classic(k, m)
        ref = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
        assert o.encode(ref) == 0
        if n % 3 == 0:
            items.append((codecs[(k, m)], rs.OP_ENCODE, shards))
            expect.append(("enc", k, m, ref))
        else:
            # reconstruct from a valid codeword with random erasures
            for i in range(k, k + m):
                shards[i].copy_(torch.from_numpy(ref[i]))
            e = int(rng.integers(0, m + 1))
            lost = set(rng.choice(k + m, e, replace=False).tolist())
            work = [None if i in lost else shards[i] for i in range(k + m)]
            op = rs.OP_RECONSTRUCT if n % 3 == 1 else rs.OP_RECONSTRUCT_DATA
            items.append((codecs[(k, m)], op, work))
            expect.append(("rec" if op == rs.OP_RECONSTRUCT else "recdata", k, m, ref, lost))
    b = rs.Batch(0)
    status = b.run(items)
    assert all(s == 0 for s in status), status
    assert b.num_tasks() > 0
    for (enc, op, shards), exp in zip(items, expect):
        kind, k, m = exp[:3]
        ref = exp[3]
        if kind == "enc":
            for i in range(k, k + m):
                assert np.array_equal(shards[i].cpu().numpy(), ref[i]), ("enc", k, m, i)
        else:
            lost = exp[4]
            for i in range(k + m):
                if i in lost and (kind == "rec" or i < k):
                    assert shards[i] is not None, (kind, i)
                    assert np.array_equal(shards[i].cpu().numpy(), ref[i]), (kind, k, m, i)
                elif i in lost:
                    assert shards[i] is None
    # a plan may be replayed: outputs are overwritten with the same bytes
    b.launch()
    b.synchronize()
    for (enc, op, shards), exp in zip(items, expect):
        if exp[0] == "enc":
            k, m, ref = exp[1], exp[2], exp[3]
            for i in range(k, k + m):
                assert np.array_equal(shards[i].cpu().numpy(), ref[i])


def test_batch_item_errors_match_plain_calls():
    rng = np.random.default_rng(7)
    enc = rs.New(5, 2)
    _, good = _stripe(rng, 5, 2, 256)
    _, bad = _stripe(rng, 5, 2, 256)
    bad[3] = bad[3][:100]
    too_few = [None, None, None, good[3], good[4], good[5], good[6]]
    b = rs.Batch(0)
    status = b.run([(enc, rs.OP_ENCODE, good), (enc, rs.OP_ENCODE, bad),
                    (enc, rs.OP_RECONSTRUCT, too_few), (enc, rs.OP_ENCODE, good[:6])])
    assert status == [0, rs.ErrShardSize.code, rs.ErrTooFewShards.code, rs.ErrTooFewShards.code]


@pytest.mark.parametrize("tile,bpc", [(4096, 1), (65536, 4), (1 << 20, 8)])
def test_manual_tuning_gives_same_bytes(tile, bpc):
    rng = np.random.default_rng(tile)
    enc = rs.New(10, 4)
    data, shards = _stripe(rng, 10, 4, 300000)
    o = rsref.Classic(10, 4)
    ref = [d.copy() for d in data] + [np.zeros(300000, np.uint8) for _ in range(4)]
    o.encode(ref)
    b = rs.Batch(0)
    b.set_tuning(tile, bpc)
    assert b.run([(enc, rs.OP_ENCODE, shards)]) == [0]
    assert b.tuning() == (tile, bpc)
    for i in range(10, 14):
        assert np.array_equal(shards[i].cpu().numpy(), ref[i])
