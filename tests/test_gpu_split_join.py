"""Split / Join (reedsolomon.go:1803-1910, leopard.go / leopard8.go twins)
through encode -> verify -> erase -> reconstruct, as the reference's
TestSplitJoin (reedsolomon_test.go:1782-1850), host and device buffers."""
import io

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import reedsolomon_amd as rs  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available() or rs.device_count() == 0:
        pytest.skip("no GPU")


OPTS = [(), (rs.WithLeopardGF(True),), (rs.WithLeopardGF16(True),)]


@pytest.mark.parametrize("mode", ["host", "device"])
@pytest.mark.parametrize("opt", range(3))
def test_split_join_round_trip(mode, opt):
    rng = np.random.default_rng(opt)
    for k, m in [(1, 0), (5, 0), (5, 1), (12, 4), (2, 15), (17, 1)]:
        enc = rs.New(k, m, *OPTS[opt])
        with pytest.raises(rs.ErrShortData):
            enc.Split(np.zeros(0, np.uint8))
        with pytest.raises(rs.ErrTooFewShards):
            enc.Join(io.BytesIO(), [], 0)
        for size in (k, 1337, 2699):
            ref = rng.integers(0, 256, size, dtype=np.uint8)
            data = ref.copy() if mode == "host" else torch.from_numpy(ref.copy()).to("cuda")
            shards = enc.Split(data)
            assert len(shards) == k + m
            per = len(shards[0])
            assert all(len(s) == per for s in shards) and per % enc.ShardSizeMultiple() == 0
            enc.Encode(shards)
            assert enc.Verify(shards)
            for i in range(m):
                shards[i] = None
            enc.Reconstruct(shards)
            buf = io.BytesIO()
            enc.Join(buf, shards, size)
            assert buf.getvalue() == ref.tobytes(), (k, m, size)
            with pytest.raises(rs.ErrShortData):
                enc.Join(io.BytesIO(), shards, per * k + 1)
            if k > 1:
                holes = list(shards)
                holes[0] = None
                with pytest.raises(rs.ErrReconstructRequired):
                    enc.Join(io.BytesIO(), holes, size)


@pytest.mark.parametrize("opt", range(3))
def test_shard_files_encode_decode(tmp_path, opt):
    # examples/simple-encoder.go and simple-decoder.go: file -> name.N shards
    # -> two shard files lost -> Verify, Reconstruct, Join -> the file again
    # (zero-padded to len(shards[0]) * DataShards, as the example writes it).
    enc = rs.New(6, 3, *OPTS[opt])
    raw = np.random.default_rng(40 + opt).integers(0, 256, 100_003, dtype=np.uint8)
    src = tmp_path / "file.bin"
    src.write_bytes(raw.tobytes())
    names = rs.encode_file(enc, str(src))
    assert len(names) == 9
    out = tmp_path / "file.out"
    assert not rs.decode_files(enc, str(src), str(out))
    got = out.read_bytes()
    assert got[:len(raw)] == raw.tobytes() and not any(got[len(raw):])
    for i in (1, 7):
        (tmp_path / f"file.bin.{i}").unlink()
    assert rs.decode_files(enc, str(src), str(out), out_size=len(raw))
    assert out.read_bytes() == raw.tobytes()
