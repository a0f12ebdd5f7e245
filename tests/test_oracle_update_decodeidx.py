"""Oracle of Update (reedsolomon.go:762-818) and DecodeIdx (reedsolomon.go:
1464-1617), pinned by what the mathematics fixes: after Update the parity is
the encode of the new data and each changed old shard holds old ^ new;
DecodeIdx of a valid codeword rebuilds the original bytes, progressively (one
input per call) and merged from partial decodings.  Error cases follow the
reference's own tests (reedsolomon_test.go testUpdate, decodeidx_test.go)."""
import numpy as np
import pytest

from oracle import rsref


def _codeword(o, k, m, size, rng):
    sh = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + [np.zeros(size, np.uint8)
                                                                           for _ in range(m)]
    assert o.encode(sh) == 0
    return sh


@pytest.mark.parametrize("k,m", [(10, 3), (17, 2), (4, 4), (1, 1), (5, 0)])
def test_update_equals_encode_of_new_data(k, m):
    rng = np.random.default_rng(k * 10 + m)
    o = rsref.Classic(k, m)
    size = 777
    shards = _codeword(o, k, m, size, rng)
    # single shards, then pairs, then windows (reedsolomon_test.go:573-640)
    patterns = [[s] for s in range(k)] + [[s, s + 1] for s in range(k - 1)] + [list(range(k))]
    for pat in patterns:
        new = [None] * k
        for s in pat:
            new[s] = rng.integers(0, 256, size, dtype=np.uint8)
        old = [x.copy() for x in shards]
        cur = [x.copy() for x in shards]
        cur_in = [cur[i] if i in pat else None for i in range(k)] + cur[k:]
        assert o.update(cur_in, new) == 0
        for s in pat:
            assert np.array_equal(cur_in[s], old[s] ^ new[s])  # old becomes the delta
        for s in pat:
            shards[s] = new[s]
        ref = [x.copy() for x in shards[:k]] + [np.zeros(size, np.uint8) for _ in range(m)]
        assert o.encode(ref) == 0
        for i in range(k, k + m):
            assert np.array_equal(cur_in[i], ref[i]), (pat, i)
        shards = shards[:k] + [cur_in[i] for i in range(k, k + m)]
        assert o.verify(shards) == (0, True)


def test_update_errors():
    k, m, size = 5, 2, 64
    rng = np.random.default_rng(1)
    o = rsref.Classic(k, m)
    sh = _codeword(o, k, m, size, rng)
    new = [None] * k
    new[1] = np.zeros(size, np.uint8)
    assert o.update(sh[:-1], new) == rsref.ERR_TOO_FEW_SHARDS
    assert o.update(sh, new[:-1]) == rsref.ERR_TOO_FEW_SHARDS
    assert o.update([None] * (k + m), new) == rsref.ERR_SHARD_NO_DATA
    bad = list(sh)
    bad[1] = None
    assert o.update(bad, new) == rsref.ERR_INVALID_INPUT       # new data for a nil old shard
    bad = list(sh)
    bad[k] = None
    assert o.update(bad, new) == rsref.ERR_INVALID_INPUT       # nil parity
    bad = list(sh)
    bad[2] = sh[2][:10]
    assert o.update(bad, new) == rsref.ERR_SHARD_SIZE
    short = [None] * k
    short[1] = np.zeros(10, np.uint8)
    assert o.update(sh, short) == rsref.ERR_SHARD_SIZE         # new data of another length


def _expect(total, marked):
    e = [False] * total
    for i in marked:
        e[i] = True
    return e


def test_decode_idx_progressive_and_merge():
    rng = np.random.default_rng(7)
    for k, m in [(4, 2), (10, 4), (5, 3), (17, 3)]:
        o = rsref.Classic(k, m)
        size = 300
        sh = _codeword(o, k, m, size, rng)
        total = k + m
        lost = sorted(rng.choice(total, m, replace=False).tolist())
        expect = [i not in lost for i in range(total)]
        # progressive: one input per call
        dst = [np.zeros(size, np.uint8) if i in lost else None for i in range(total)]
        for j in range(total):
            if expect[j]:
                inp = [None] * total
                inp[j] = sh[j]
                assert o.decode_idx(dst, expect, inp) == 0
        for i in lost:
            assert np.array_equal(dst[i], sh[i]), (k, m, i)
        # two partial decodings merged (decodeidx_test.go MergeTwoPartialDecodings)
        valid = [i for i in range(total) if expect[i]]
        half = valid[: len(valid) // 2]
        d1 = [np.zeros(size, np.uint8) if i in lost else None for i in range(total)]
        d2 = [np.zeros(size, np.uint8) if i in lost else None for i in range(total)]
        assert o.decode_idx(d1, expect, [sh[i] if i in half else None for i in range(total)]) == 0
        assert o.decode_idx(d2, expect, [sh[i] if i in valid and i not in half else None
                                         for i in range(total)]) == 0
        assert o.decode_idx(d1, None, d2) == 0
        for i in lost:
            assert np.array_equal(d1[i], sh[i])


def test_decode_idx_excess_valid_shards():
    # decodeidx_test.go:742-800: the matrix uses the first k marked shards;
    # inputs past them add nothing.
    rng = np.random.default_rng(3)
    k, m, size = 5, 3, 256
    o = rsref.Classic(k, m)
    sh = _codeword(o, k, m, size, rng)
    expect = _expect(k + m, [0, 2, 3, 4, 5, 6, 7])
    dst = [None] * (k + m)
    dst[1] = np.zeros(size, np.uint8)
    assert o.decode_idx(dst, expect, [sh[i] if i in (0, 2, 3, 4, 5) else None for i in range(k + m)]) == 0
    assert np.array_equal(dst[1], sh[1])
    dst2 = [None] * (k + m)
    dst2[1] = np.zeros(size, np.uint8)
    assert o.decode_idx(dst2, expect, [sh[i] if i in (6, 7) else None for i in range(k + m)]) == 0
    assert not dst2[1].any()


def test_decode_idx_errors():
    k, m, size = 5, 3, 64
    rng = np.random.default_rng(4)
    o = rsref.Classic(k, m)
    sh = _codeword(o, k, m, size, rng)
    total = k + m
    expect = _expect(total, range(k))
    dst = [None] * total
    dst[k] = np.zeros(size, np.uint8)
    inp = [sh[i] if i < k else None for i in range(total)]
    assert o.decode_idx(dst, expect[:-1], inp) == rsref.ERR_INVALID_INPUT       # InvalidExpectInputLength
    assert o.decode_idx(dst[:-1], expect, inp) == rsref.ERR_INVALID_INPUT       # WrongSliceLengths
    assert o.decode_idx(dst, expect, inp[:-1]) == rsref.ERR_INVALID_INPUT
    filled = list(dst)
    filled[0] = np.zeros(size, np.uint8)
    assert o.decode_idx(filled, expect, inp) == rsref.ERR_INVALID_INPUT         # DestinationAlreadyFilled
    unexpected = list(inp)
    unexpected[k + 1] = sh[k + 1]
    assert o.decode_idx(dst, expect, unexpected) == rsref.ERR_INVALID_INPUT     # UnexpectedInput
    assert o.decode_idx(dst, _expect(total, range(k - 1)),
                        [sh[i] if i < k - 1 else None for i in range(total)]) == rsref.ERR_TOO_FEW_SHARDS
    mism = list(inp)
    mism[2] = sh[2][:32]
    assert o.decode_idx(dst, expect, mism) == rsref.ERR_INVALID_SHARD_SIZE      # MismatchedBufferSizes
    small = list(dst)
    small[k] = np.zeros(32, np.uint8)
    assert o.decode_idx(small, expect, inp) == rsref.ERR_INVALID_SHARD_SIZE
    # merge mode
    assert o.decode_idx(dst, None, inp[:-1]) == rsref.ERR_INVALID_INPUT
    a = [np.zeros(size, np.uint8), None]
    b = [np.ones(size, np.uint8), np.ones(size, np.uint8)]
    assert o.decode_idx(a, None, b) == rsref.ERR_INVALID_INPUT  # input without dst
    assert a[0].all()  # the pair before the bad one was merged, as in the reference
    assert o.decode_idx([np.zeros(size, np.uint8)], None, [np.zeros(10, np.uint8)]) == rsref.ERR_INVALID_SHARD_SIZE
