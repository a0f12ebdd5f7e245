"""GPU parity of the LowLevel helpers (mulslice.go, gf16.go) with the oracle,
device and host memory, aligned and unaligned buffers."""
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


def _dev(a, offset=0):
    buf = torch.zeros(len(a) + offset + 16, dtype=torch.uint8, device="cuda")
    t = buf[offset:offset + len(a)]
    t.copy_(torch.from_numpy(a))
    return t


@pytest.mark.parametrize("mode", ["device", "device+1", "host"])
@pytest.mark.parametrize("size", [1, 63, 4096, 100003])
def test_gal_mul_slice(mode, size):
    rng = np.random.default_rng(size)
    ll = rs.LowLevel(0)
    inp = rng.integers(0, 256, size, dtype=np.uint8)
    base = rng.integers(0, 256, size, dtype=np.uint8)
    for c in (0, 1, 2, 29, 177, 255):
        for xor in (False, True):
            want = base.copy()
            rsref.gal_mul_slice(c, inp, want, xor)
            if mode == "host":
                got = base.copy()
                (ll.GalMulSliceXor if xor else ll.GalMulSlice)(c, inp.copy(), got)
            else:
                off = 1 if mode == "device+1" else 0
                got_t = _dev(base, off)
                (ll.GalMulSliceXor if xor else ll.GalMulSlice)(c, _dev(inp, off), got_t)
                got = got_t.cpu().numpy()
            assert np.array_equal(got, want), (c, xor)


def test_scalar_helpers():
    g = rsref.gf8_tables()
    for e in range(256):
        assert rs.Inv(e) == int(g["inv"][e])
    rng = np.random.default_rng(1)
    ll = rs.LowLevel(0)
    for a, b in rng.integers(0, 65536, (200, 2)):
        assert ll.GF16Mul(int(a), int(b)) == rsref.gf16_mul(int(a), int(b))


@pytest.mark.parametrize("mode", ["device", "device+16", "device+1", "host"])
def test_gf16_mul_slice_xor(mode):
    rng = np.random.default_rng(16)
    ll = rs.LowLevel(0)
    size = 64 * 1025
    inp = rng.integers(0, 256, size, dtype=np.uint8)
    outs0 = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(8)]
    scalars = [int(x) for x in rng.integers(1, 65536, 8)]
    scalars[3] = 0  # a zero scalar contributes nothing
    want = [o.copy() for o in outs0]
    for s, w in zip(scalars, want):
        rsref.gf16_mul_slice_xor(s, inp, w)
    if mode == "host":
        got = [o.copy() for o in outs0]
        ll.GF16MulSliceXor8(scalars, inp.copy(), got)
        single = outs0[0].copy()
        ll.GF16MulSliceXor(scalars[0], inp.copy(), single)
    else:
        off = {"device": 0, "device+16": 16, "device+1": 1}[mode]
        tin = _dev(inp, off)
        tg = [_dev(o, off) for o in outs0]
        ll.GF16MulSliceXor8(scalars, tin, tg)
        got = [t.cpu().numpy() for t in tg]
        ts = _dev(outs0[0], off)
        ll.GF16MulSliceXor(scalars[0], tin, ts)
        single = ts.cpu().numpy()
    for j in range(8):
        assert np.array_equal(got[j], want[j]), j
    assert np.array_equal(single, want[0])


def test_gf16_mul_slice_xor_errors():
    ll = rs.LowLevel(0)
    a = torch.zeros(100, dtype=torch.uint8, device="cuda")
    with pytest.raises(rs.ErrInvalidShardSize):
        ll.GF16MulSliceXor(3, a, torch.zeros(100, dtype=torch.uint8, device="cuda"))
    with pytest.raises(rs.ErrInvalidShardSize):
        ll.GF16MulSliceXor(3, torch.zeros(128, dtype=torch.uint8, device="cuda"),
                           torch.zeros(64, dtype=torch.uint8, device="cuda"))
    ll.GF16MulSliceXor(3, torch.zeros(0, dtype=torch.uint8, device="cuda"), a)  # empty input: no-op
