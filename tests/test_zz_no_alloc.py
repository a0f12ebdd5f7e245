"""rsx_reconstruct_noalloc (the allocation-free Reconstruct / ReconstructData
of the classic engine) against rsx_reconstruct / rsx_reconstruct_data,
bit for bit, through the C ABI: two calls in a row, the second on smaller
shards, then other shapes sharing the per-GPU scratch array, and the
rejections (tables beyond the scratch, Leopard codecs, host memory)."""
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


def _codeword(k, m, size, seed):
    rng = np.random.default_rng(seed)
    shards = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + \
             [np.zeros(size, np.uint8) for _ in range(m)]
    assert rsref.Classic(k, m).encode(shards) == 0
    return shards


def _dev(full, lost):
    return [None if i in lost else torch.from_numpy(full[i].copy()).to("cuda") for i in range(len(full))]


def _check_pair(enc, full, lost, data_only):
    k = enc.DataShards()
    base = _dev(full, lost)
    (enc.ReconstructData if data_only else enc.Reconstruct)(base)
    work = _dev(full, lost)
    enc.ReconstructNoAlloc(work, data_only=data_only)
    for i in range(len(full)):
        if base[i] is None:
            assert work[i] is None, i
            assert data_only and i >= k
            continue
        got = work[i].cpu().numpy()
        assert np.array_equal(got, base[i].cpu().numpy()), (i, lost)
        assert np.array_equal(got, full[i]), (i, lost)


def test_two_calls_second_smaller():
    enc = rs.New(10, 4)
    big = _codeword(10, 4, 1 << 20, 1)
    _check_pair(enc, big, {1, 5, 11, 13}, data_only=False)
    small = _codeword(10, 4, 4099, 2)
    _check_pair(enc, small, {0, 12}, data_only=True)


@pytest.mark.parametrize("k,m,lost,size", [
    (17, 3, {2, 9, 18}, 65536),
    (200, 16, set(range(0, 32, 2)), 8192),  # 3200 coefficients: two table launches
    (32, 8, {0, 1, 2, 3, 36, 37, 38, 39}, 100003),
    (4, 2, {4}, 64),
])
def test_shapes_share_the_scratch(k, m, lost, size):
    enc = rs.New(k, m)
    full = _codeword(k, m, size, k * 1000 + m)
    _check_pair(enc, full, lost, data_only=False)
    _check_pair(enc, full, lost, data_only=True)


def test_rejections():
    enc = rs.New(128, 128)
    full = _codeword(128, 128, 64, 3)
    with pytest.raises(rs.ErrNotSupported):  # 128 x 64 coefficients > 3276
        enc.ReconstructNoAlloc(_dev(full, set(range(64))))
    leo = rs.New(8, 4, rs.WithLeopardGF16(True))
    shards = [torch.zeros(64, dtype=torch.uint8, device="cuda") for _ in range(12)]
    shards[0] = None
    with pytest.raises(rs.ErrNotSupported):
        leo.ReconstructNoAlloc(shards)
    enc = rs.New(4, 2)
    host = _codeword(4, 2, 256, 4)
    host[1] = None
    with pytest.raises(rs.ErrMixedMemory):
        enc.ReconstructNoAlloc(host)
