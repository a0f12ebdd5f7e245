"""Oracle of the LowLevel helpers (mulslice.go, gf16.go): GF(2^16) products
pinned by gf16_test.go's field properties and by agreement of the slice
multiply with the scalar GF16Mul restatement; GF(2^8) slice products by the
galois_test.go KATs (test_oracle_classic)."""
import numpy as np

from oracle import rsref


def test_gf16_mul_field_properties():
    # gf16_test.go: zero, one, commutativity, associativity, distributivity.
    for v in (0, 1, 2, 100, 65535):
        assert rsref.gf16_mul(v, 0) == 0 and rsref.gf16_mul(0, v) == 0
    for v in (1, 2, 7, 255, 1000, 65535):
        assert rsref.gf16_mul(v, 1) == v and rsref.gf16_mul(1, v) == v
    for a, b in ((1, 2), (3, 5), (255, 256), (1000, 2000), (65534, 65535)):
        assert rsref.gf16_mul(a, b) == rsref.gf16_mul(b, a)
    assert rsref.gf16_mul(rsref.gf16_mul(3, 5), 7) == rsref.gf16_mul(3, rsref.gf16_mul(5, 7))
    for a, b, c in ((3, 5, 7), (255, 256, 1000), (1, 2, 3), (65534, 65535, 1)):
        assert rsref.gf16_mul(a, b ^ c) == rsref.gf16_mul(a, b) ^ rsref.gf16_mul(a, c)
    # every non-zero element has an inverse (a field, not just a ring)
    rng = np.random.default_rng(0)
    for a in rng.integers(1, 65536, 20):
        assert any(rsref.gf16_mul(int(a), int(b)) == 1 for b in range(1, 65536))


def test_gf16_mul_slice_xor_matches_scalar_products():
    rng = np.random.default_rng(5)
    inp = rng.integers(0, 256, 64 * 3, dtype=np.uint8)
    out0 = rng.integers(0, 256, 64 * 3, dtype=np.uint8)
    for scalar in (1, 2, 3, 1000, 65535, 0):
        out = out0.copy()
        rsref.gf16_mul_slice_xor(scalar, inp, out)
        c, o = inp.reshape(-1, 2, 32), out0.reshape(-1, 2, 32)
        got = out.reshape(-1, 2, 32)
        for ch in range(c.shape[0]):
            for i in range(32):
                x = int(c[ch, 0, i]) | (int(c[ch, 1, i]) << 8)
                want = (int(o[ch, 0, i]) | (int(o[ch, 1, i]) << 8)) ^ rsref.gf16_mul(scalar, x)
                assert int(got[ch, 0, i]) | (int(got[ch, 1, i]) << 8) == want
