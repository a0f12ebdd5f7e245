"""GPU StreamEncoder (streaming.go) against the in-memory oracle: parity of
every block, Verify, Reconstruct of data and parity, Split/Join, and the
reference's error behaviour (streaming_test.go)."""
import io

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


def _encode_ref(k, m, data):
    size = len(data[0])
    sh = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
    assert rsref.Classic(k, m).encode(sh) == 0
    return sh


class _Fail:
    def read(self, n):
        raise IOError("boom")

    def write(self, b):
        raise IOError("boom")


@pytest.mark.parametrize("k,m,size,bs", [(10, 4, 10 << 20, 4 << 20), (5, 3, 1000, 256), (3, 2, 4096, 4096),
                                         (17, 3, 123457, 65536)])
def test_stream_encode_verify_reconstruct(k, m, size, bs):
    rng = np.random.default_rng(size + k)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    ref = _encode_ref(k, m, data)  # column-separable: blocks encode like the whole
    enc = rs.NewStream(k, m, block_size=bs)
    par = [io.BytesIO() for _ in range(m)]
    enc.Encode([io.BytesIO(d.tobytes()) for d in data], par)
    for i in range(m):
        assert par[i].getvalue() == ref[k + i].tobytes(), i
    assert enc.Verify([io.BytesIO(x.tobytes()) for x in ref])
    bad = [x.copy() for x in ref]
    bad[1][size // 2] ^= 1
    assert not enc.Verify([io.BytesIO(x.tobytes()) for x in bad])
    # reconstruct a data shard and a parity shard
    lost = [0, k]
    valid = [None if i in lost else io.BytesIO(ref[i].tobytes()) for i in range(k + m)]
    fill = [io.BytesIO() if i in lost else None for i in range(k + m)]
    enc.Reconstruct(valid, fill)
    for i in lost:
        assert fill[i].getvalue() == ref[i].tobytes(), i
    # data only
    valid = [None if i == 1 else io.BytesIO(ref[i].tobytes()) for i in range(k + m)]
    fill = [io.BytesIO() if i == 1 else None for i in range(k + m)]
    enc.Reconstruct(valid, fill)
    assert fill[1].getvalue() == ref[1].tobytes()


def test_stream_errors():
    k, m = 5, 3
    enc = rs.NewStream(k, m, block_size=1024)
    rng = np.random.default_rng(2)
    data = [rng.integers(0, 256, 3000, dtype=np.uint8) for _ in range(k)]
    with pytest.raises(rs.ErrTooFewShards):
        enc.Encode([io.BytesIO(d.tobytes()) for d in data[:-1]], [io.BytesIO() for _ in range(m)])
    with pytest.raises(rs.ErrShardNoData):
        enc.Encode([io.BytesIO() for _ in range(k)], [io.BytesIO() for _ in range(m)])
    with pytest.raises(rs.ErrShardSize):
        enc.Encode([io.BytesIO(d.tobytes()) for d in data[:-1]] + [io.BytesIO(data[-1][:2500].tobytes())],
                   [io.BytesIO() for _ in range(m)])
    with pytest.raises(rs.StreamReadError) as ei:
        enc.Encode([io.BytesIO(d.tobytes()) for d in data[:2]] + [_Fail()] + [io.BytesIO(d.tobytes())
                                                                              for d in data[3:]],
                   [io.BytesIO() for _ in range(m)])
    assert ei.value.stream == 2
    with pytest.raises(rs.StreamWriteError) as ei:
        enc.Encode([io.BytesIO(d.tobytes()) for d in data], [io.BytesIO(), _Fail(), io.BytesIO()])
    assert ei.value.stream == 1
    with pytest.raises(rs.ErrReconstructMismatch):
        enc.Reconstruct([io.BytesIO()] * (k + m), [io.BytesIO()] + [None] * (k + m - 1))
    with pytest.raises(rs.ErrTooFewShards):
        enc.Reconstruct([None] * 4 + [io.BytesIO(d.tobytes()) for d in data[:1]] + [None] * 3,
                        [io.BytesIO()] * 4 + [None] * 4)


@pytest.mark.parametrize("size", [1, 1337, 50000])
def test_stream_split_join(size):
    k, m = 5, 3
    enc = rs.NewStream(k, m, block_size=4096)
    rng = np.random.default_rng(size)
    raw = rng.integers(0, 256, size, dtype=np.uint8).tobytes()
    dst = [io.BytesIO() for _ in range(k)]
    enc.Split(io.BytesIO(raw), dst, size)
    per = (size + k - 1) // k
    assert all(len(d.getvalue()) == per for d in dst)
    shards = [np.frombuffer(d.getvalue(), np.uint8).copy() for d in dst]
    par = [io.BytesIO() for _ in range(m)]
    enc.Encode([io.BytesIO(s.tobytes()) for s in shards], par)
    out = io.BytesIO()
    enc.Join(out, [io.BytesIO(s.tobytes()) for s in shards], size)
    assert out.getvalue() == raw
    with pytest.raises(rs.ErrShortData):
        enc.Join(io.BytesIO(), [io.BytesIO(s.tobytes()) for s in shards], per * k + 1)
    with pytest.raises(rs.ErrShortData):
        enc.Split(io.BytesIO(raw), [io.BytesIO() for _ in range(k)], 0)


@pytest.mark.parametrize("reads,writes", [(True, True), (True, False), (False, True)])
def test_concurrent_streams_match_sequential(reads, writes):
    # streaming.go:334-420 cReadShards / cWriteShards (NewStreamC): the same
    # bytes as the sequential streams; a failing reader or writer is a
    # StreamReadError / StreamWriteError naming a failing stream (here the
    # lowest index of the failing ones).
    k, m, size = 12, 4, 300001
    rng = np.random.default_rng(77)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    ref = _encode_ref(k, m, data)
    enc = rs.NewStreamC(k, m, reads, writes, rs.WithStreamBlockSize(65536))
    par = [io.BytesIO() for _ in range(m)]
    enc.Encode([io.BytesIO(d.tobytes()) for d in data], par)
    assert all(par[i].getvalue() == ref[k + i].tobytes() for i in range(m))
    assert enc.Verify([io.BytesIO(x.tobytes()) for x in ref])
    valid = [None if i in (2, k + 1) else io.BytesIO(ref[i].tobytes()) for i in range(k + m)]
    fill = [io.BytesIO() if i in (2, k + 1) else None for i in range(k + m)]
    enc.Reconstruct(valid, fill)
    assert fill[2].getvalue() == ref[2].tobytes() and fill[k + 1].getvalue() == ref[k + 1].tobytes()
    if reads:
        bad = [io.BytesIO(d.tobytes()) for d in data]
        bad[7], bad[3] = _Fail(), _Fail()
        with pytest.raises(rs.StreamReadError) as e:
            enc.Encode(bad, [io.BytesIO() for _ in range(m)])
        assert e.value.stream == 3
    if writes:
        outs = [io.BytesIO() for _ in range(m)]
        outs[2] = _Fail()
        with pytest.raises(rs.StreamWriteError) as e:
            enc.Encode([io.BytesIO(d.tobytes()) for d in data], outs)
        assert e.value.stream == 2


def test_stream_block_size_options():
    # WithStreamBlockSize, else a positive WithAutoGoroutines shard size, sets
    # the block (streaming.go:157-163); any block size gives the same parity.
    k, m, size = 6, 3, 100000
    rng = np.random.default_rng(5)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    ref = _encode_ref(k, m, data)
    for opts in ((rs.WithStreamBlockSize(4096),), (rs.WithAutoGoroutines(10000),),
                 (rs.WithAutoGoroutines(0), rs.WithStreamBlockSize(777))):
        enc = rs.NewStream(k, m, *opts)
        par = [io.BytesIO() for _ in range(m)]
        enc.Encode([io.BytesIO(d.tobytes()) for d in data], par)
        assert all(par[i].getvalue() == ref[k + i].tobytes() for i in range(m)), opts
