"""GPU parity of Update and DecodeIdx with the oracle: every byte of the
parity, of the changed old shards (old ^ new) and of the decoded outputs, and
every error code, for shards in device and in host memory."""
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


def _to(mode, x):
    if x is None:
        return None
    return torch.from_numpy(x.copy()).to("cuda") if mode == "device" else x.copy()


def _np(x):
    return None if x is None else (x.cpu().numpy() if torch.is_tensor(x) else x)


def _code(fn):
    try:
        fn()
        return 0
    except rs.RSError as e:
        return e.code


@pytest.mark.parametrize("mode", ["device", "host"])
@pytest.mark.parametrize("k,m,size", [(10, 3, 4096), (17, 2, 1000), (32, 8, 65536 + 13), (4, 4, 1), (5, 0, 100)])
def test_update_matches_oracle(mode, k, m, size):
    rng = np.random.default_rng(k * 1000 + m + size)
    o = rsref.Classic(k, m)
    enc = rs.New(k, m)
    sh = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + [np.zeros(size, np.uint8)
                                                                           for _ in range(m)]
    assert o.encode(sh) == 0
    for trial in range(4):
        changed = sorted(rng.choice(k, int(rng.integers(1, k + 1)), replace=False).tolist())
        new = [rng.integers(0, 256, size, dtype=np.uint8) if i in changed else None for i in range(k)]
        want = [x.copy() if (i >= k or i in changed) else None for i, x in enumerate(sh)]
        assert o.update(want, [None if x is None else x.copy() for x in new]) == 0
        got = [_to(mode, x) if (i >= k or i in changed) else None for i, x in enumerate(sh)]
        enc.Update(got, [_to(mode, x) for x in new])
        for i in range(k + m):
            if want[i] is not None:
                assert np.array_equal(_np(got[i]), want[i]), (mode, trial, i)
        for i in changed:
            sh[i] = new[i]
        sh = sh[:k] + [want[i] for i in range(k, k + m)]


@pytest.mark.parametrize("mode", ["device", "host"])
def test_update_errors_match_oracle(mode):
    rng = np.random.default_rng(11)
    k, m, size = 5, 2, 64
    o = rsref.Classic(k, m)
    enc = rs.New(k, m)
    sh = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k + m)]
    new = [None, np.zeros(size, np.uint8), None, None, None]
    cases = [(sh[:-1], new), (sh, new[:-1]), ([None] * (k + m), new),
             ([sh[0], None] + sh[2:], new), (sh[:k] + [None, sh[k + 1]], new),
             (sh[:2] + [sh[2][:10]] + sh[3:], new), (sh, [None, np.zeros(10, np.uint8), None, None, None])]
    for shards, nd in cases:
        want = o.update([None if x is None else x.copy() for x in shards], [None if x is None else x.copy() for x in nd])
        got = _code(lambda: enc.Update([_to(mode, x) for x in shards], [_to(mode, x) for x in nd]))
        assert got == want, (len(shards), len(nd), want, got)
    leo = rs.New(4, 2, rs.WithLeopardGF16(True))
    assert _code(lambda: leo.Update([_to(mode, x[:64]) for x in sh[:6]], [None] * 4)) == rs.ErrNotSupported.code


@pytest.mark.parametrize("mode", ["device", "host"])
@pytest.mark.parametrize("k,m,size", [(4, 2, 300), (10, 4, 65536), (17, 3, 4097), (32, 8, 1000)])
def test_decode_idx_matches_oracle(mode, k, m, size):
    rng = np.random.default_rng(k * 7 + m + size)
    o = rsref.Classic(k, m)
    enc = rs.New(k, m)
    total = k + m
    sh = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + [np.zeros(size, np.uint8)
                                                                           for _ in range(m)]
    assert o.encode(sh) == 0
    for trial in range(3):
        lost = sorted(rng.choice(total, int(rng.integers(1, m + 1)), replace=False).tolist())
        extra = [i for i in range(total) if i not in lost]
        expect = [i not in lost for i in range(total)]
        # inputs arrive in random groups, one DecodeIdx call per group
        order = rng.permutation(extra).tolist()
        groups = np.array_split(order, int(rng.integers(1, len(order) + 1)))
        want = [np.zeros(size, np.uint8) if i in lost else None for i in range(total)]
        got = [_to(mode, x) for x in want]
        for g in groups:
            inp = [sh[i] if i in g else None for i in range(total)]
            assert o.decode_idx(want, expect, inp) == 0
            enc.DecodeIdx(got, expect, [_to(mode, x) for x in inp])
        for i in lost:
            assert np.array_equal(want[i], sh[i])
            assert np.array_equal(_np(got[i]), want[i]), (mode, trial, i)
    # merge mode: two partial results combined
    lost = list(range(m))
    expect = [i not in lost for i in range(total)]
    valid = [i for i in range(total) if expect[i]]
    d1 = [_to(mode, np.zeros(size, np.uint8)) if i in lost else None for i in range(total)]
    d2 = [_to(mode, np.zeros(size, np.uint8)) if i in lost else None for i in range(total)]
    enc.DecodeIdx(d1, expect, [_to(mode, sh[i]) if i in valid[::2] else None for i in range(total)])
    enc.DecodeIdx(d2, expect, [_to(mode, sh[i]) if i in valid[1::2] else None for i in range(total)])
    enc.DecodeIdx(d1, None, d2)
    for i in lost:
        assert np.array_equal(_np(d1[i]), sh[i])


@pytest.mark.parametrize("mode", ["device", "host"])
def test_decode_idx_errors_match_oracle(mode):
    k, m, size = 5, 3, 64
    rng = np.random.default_rng(4)
    o = rsref.Classic(k, m)
    enc = rs.New(k, m)
    total = k + m
    sh = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(total)]
    expect = [i < k for i in range(total)]
    dst = [None] * total
    dst[k] = np.zeros(size, np.uint8)
    inp = [sh[i] if i < k else None for i in range(total)]
    filled = list(dst)
    filled[0] = np.zeros(size, np.uint8)
    unexpected = list(inp)
    unexpected[k + 1] = sh[k + 1]
    mism = list(inp)
    mism[2] = sh[2][:32]
    small = list(dst)
    small[k] = np.zeros(32, np.uint8)
    cases = [(dst, expect[:-1], inp), (dst[:-1], expect, inp), (dst, expect, inp[:-1]), (filled, expect, inp),
             (dst, expect, unexpected), (dst, [i < k - 1 for i in range(total)], [x if i < k - 1 else None
                                                                                  for i, x in enumerate(inp)]),
             (dst, expect, mism), (small, expect, inp), (dst, None, inp[:-1]),
             ([np.zeros(size, np.uint8), None], None, [np.ones(size, np.uint8)] * 2),
             ([np.zeros(size, np.uint8)], None, [np.zeros(10, np.uint8)])]
    for d, e, i in cases:
        dw = [None if x is None else x.copy() for x in d]
        want = o.decode_idx(dw, e, [None if x is None else x.copy() for x in i])
        dg = [_to(mode, x) for x in d]
        got = _code(lambda: enc.DecodeIdx(dg, e, [_to(mode, x) for x in i]))
        assert got == want, (want, got)
        for a, b in zip(dw, dg):  # partial merges match too
            if a is not None:
                assert np.array_equal(_np(b), a)
