"""Split / Join layout, host side (no GPU): the shard sizes and byte
placement of Encoder.Split (reedsolomon.go:1803-1860; leopard.go:320-376 and
leopard8.go:293-350 round the shard size up to 64) and the error order of
Encoder.Join (reedsolomon.go:1875-1910), over the shapes and sizes of
TestSplitJoin (reedsolomon_test.go:1782-1850) plus the boundary sizes
around each rounding step.

Each expectation is stated from the layout itself (shard i holds bytes
[i*per, (i+1)*per) of the input, zero beyond its end) rather than from the
functions' own arithmetic, and the per-shard size is checked against the
smallest size that satisfies the layout's two constraints (k shards hold the
input, the size is a multiple of the codec's)."""
import io
import os

import numpy as np
import pytest

import reedsolomon_amd as rs

SHAPES = [(1, 0), (5, 0), (5, 1), (12, 4), (2, 15), (17, 1), (1, 3)]


def _sizes(k, mult):
    out = {1, k, 1337, 2699, k * mult, k * mult + 1, k * mult - 1, 64 * k * 3 + 7}
    return sorted(s for s in out if s > 0)


def _smallest_per(n, k, mult):
    # brute force: the least multiple of `mult` with k * per >= n
    per = mult
    while k * per < n:
        per += mult
    return per


@pytest.mark.parametrize("mult", [1, 64])
def test_split_layout(mult):
    rng = np.random.default_rng(mult)
    for k, m in SHAPES:
        total = k + m
        for n in _sizes(k, mult):
            data = rng.integers(0, 256, n, dtype=np.uint8)
            keep = data.copy()
            shards = rs.split_shards(data, k, total, mult)
            assert len(shards) == total
            if total == 1 and n % mult == 0:
                # the input comes back whole (reedsolomon.go:1807, leopard.go:324)
                assert shards[0] is data
                continue
            per = _smallest_per(n, k, mult)
            assert [len(s) for s in shards] == [per] * total, (k, m, n)
            flat = np.concatenate(shards)
            assert np.array_equal(flat[:n], keep)
            assert not flat[n:].any(), "padding and parity shards start zeroed"
            for i, s in enumerate(shards):
                whole = (i + 1) * per <= n
                assert np.shares_memory(s, data) == whole, (k, m, n, i)
            # the input is never written: padding goes into new buffers
            assert np.array_equal(data, keep)


def test_split_torch_views():
    torch = pytest.importorskip("torch")
    data = torch.arange(1000, dtype=torch.int32).to(torch.uint8)
    shards = rs.split_shards(data, 3, 5, 64)
    assert [len(s) for s in shards] == [384] * 5
    assert shards[0].data_ptr() == data.data_ptr()
    flat = torch.cat(shards)
    assert torch.equal(flat[:1000], data) and not flat[1000:].any()


def test_split_empty():
    for mult in (1, 64):
        with pytest.raises(rs.ErrShortData):
            rs.split_shards(np.zeros(0, np.uint8), 4, 6, mult)


def test_join_writes_prefix_and_errors():
    rng = np.random.default_rng(7)
    k, m, n = 5, 2, 2699
    data = rng.integers(0, 256, n, dtype=np.uint8)
    shards = rs.split_shards(data.copy(), k, k + m, 64)
    per = len(shards[0])
    for size in (0, 1, per, per + 1, n, per * k):
        buf = io.BytesIO()
        rs.join_shards(buf, shards, k, size)
        assert buf.getvalue() == np.concatenate(shards[:k]).tobytes()[:size]
    with pytest.raises(rs.ErrShortData):
        rs.join_shards(io.BytesIO(), shards, k, per * k + 1)
    with pytest.raises(rs.ErrTooFewShards):
        rs.join_shards(io.BytesIO(), shards[:k - 1], k, 1)
    holes = list(shards)
    holes[2] = None
    # a missing shard past out_size is not needed (reedsolomon.go:1893-1897)
    buf = io.BytesIO()
    rs.join_shards(buf, holes, k, 2 * per)
    assert buf.getvalue() == data[:2 * per].tobytes()
    with pytest.raises(rs.ErrReconstructRequired):
        rs.join_shards(io.BytesIO(), holes, k, 2 * per + 1)
    # parity shards are never written
    buf = io.BytesIO()
    rs.join_shards(buf, shards[:k] + [None] * m, k, n)
    assert buf.getvalue() == data.tobytes()


def test_shard_files_round_trip(tmp_path):
    data = np.arange(3000, dtype=np.uint32).astype(np.uint8)
    src = tmp_path / "payload.bin"
    src.write_bytes(data.tobytes())
    shards = rs.split_shards(data, 4, 6, 1)
    names = rs.write_shard_files(str(src), shards)
    assert names == [str(tmp_path / f"payload.bin.{i}") for i in range(6)]
    os.remove(names[1])
    back = rs.read_shard_files(str(src), 6)
    assert back[1] is None
    for i in (0, 2, 3, 4, 5):
        assert np.array_equal(back[i], shards[i])
    out = tmp_path / "elsewhere"
    out.mkdir()
    assert rs.write_shard_files(str(src), shards, out_dir=str(out))[0] == str(out / "payload.bin.0")
