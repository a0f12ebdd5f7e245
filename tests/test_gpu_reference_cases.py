"""GPU runs of reference test cases not covered elsewhere: PAR1 singular
reconstruct, a custom non-Vandermonde matrix, a Leopard shape with more
erasures than parity, Split of a zero-parity codec, and a zero-parity stream
(reedsolomon_test.go, leopard_test.go, streaming_test.go)."""
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


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def test_reconstruct_par1_singular():
    # reedsolomon_test.go:962-998: the PAR1 matrix is singular for this
    # survivor set (last data shard, parity 0, 1, 3).
    rng = np.random.default_rng(962)
    enc = rs.New(4, 4, rs.WithPAR1Matrix())
    shards = [dev(rng.integers(0, 256, 50, dtype=np.uint8)) for _ in range(8)]
    enc.Encode(shards)
    for i in (0, 1, 2, 6):
        shards[i] = None
    with pytest.raises(rs.ErrSingular):
        enc.Reconstruct(shards)


def test_reconstruct_custom_matrix():
    # reedsolomon_test.go:735-794: a custom matrix with zero coefficients;
    # one and then three data shards absent both reconstruct.
    custom = np.array([[1, 1, 0, 0], [0, 0, 1, 1], [1, 2, 3, 4]], dtype=np.uint8)
    rng = np.random.default_rng(735)
    data = [rng.integers(0, 256, 50000, dtype=np.uint8) for _ in range(4)]
    want = data + [np.zeros(50000, np.uint8) for _ in range(3)]
    assert rsref.Classic(4, 3, rsref.MATRIX_DEFAULT, custom).encode(want) == 0
    enc = rs.New(4, 3, rs.WithCustomMatrix(custom))
    shards = [dev(d) for d in data] + [torch.zeros(50000, dtype=torch.uint8, device="cuda") for _ in range(3)]
    enc.Encode(shards)
    for i in range(7):
        assert np.array_equal(shards[i].cpu().numpy(), want[i]), i
    for lost in ((0,), (0, 1, 2)):
        work = [None if i in lost else shards[i].clone() for i in range(7)]
        enc.Reconstruct(work)
        for i in lost:
            assert np.array_equal(work[i].cpu().numpy(), want[i]), (lost, i)
        assert enc.Verify(shards)


def test_leopard_more_erasures_than_parity():
    # leopard_test.go:89-124: 500 + 300 shards (GF(2^16) Leopard by shard
    # count) from 2 MiB; Verify holds; with 301 shards gone Reconstruct
    # reports ErrTooFewShards.
    rng = np.random.default_rng(89)
    enc = rs.New(500, 300)
    shards = [dev(s) for s in enc.Split(rng.integers(0, 256, 2 << 20, dtype=np.uint8))]
    enc.Encode(shards)
    assert enc.Verify(shards)
    for i in range(301):
        shards[i] = None
    with pytest.raises(rs.ErrTooFewShards):
        enc.Reconstruct(shards)


def test_split_zero_parity():
    # reedsolomon_test.go:2097-2108: Split works for one data shard and no
    # parity.
    enc = rs.New(1, 0)
    shards = enc.Split(np.zeros(512, np.uint8))
    assert len(shards) == 1 and len(shards[0]) >= 512


def test_stream_zero_parity():
    # streaming_test.go:119-156: Encode with no parity writers, Verify true,
    # Reconstruct with nothing to fill.
    rng = np.random.default_rng(119)
    data = [rng.integers(0, 256, 50000, dtype=np.uint8).tobytes() for _ in range(10)]
    enc = rs.NewStream(10, 0)
    enc.Encode([io.BytesIO(d) for d in data], [])
    assert enc.Verify([io.BytesIO(d) for d in data])
    enc.Reconstruct([io.BytesIO(d) for d in data], [None] * 10)
