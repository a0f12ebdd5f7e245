"""Pins the Leopard GF(2^16) / GF(2^8) field representation, the FFT skews,
logWalsh and the encoder's parity bytes to the reference, for the oracle and
for the engine's own tables (read through rsx_leopard_tables / rsx_gf16_mul,
host functions that need no GPU).

The reference fixes the bytes only through its constants: the polynomial and
the Cantor basis literals (leopard.go:100-105, 1105-1110; leopard8.go:96-104,
1030-1032), parsed from the source text into tests/golden/leopard_consts.json.
Everything else is checked against mathematics that does not share a step
with the restatements:

  * a field element's bits are its coordinates over the Cantor basis
    (leopard.go:1127-1139), so phi(a) = XOR of basis[i] over the set bits of a
    maps it into GF(2)[x]/polynomial, where a product is a carry-less
    multiply and a reduction.  gf16_mul / expLUT / logLUT are checked against
    that, with no log or exp table involved;
  * fftSkew (leopard.go:1150-1195) is the log of the normalised subspace
    polynomial W_m(x) = prod_{a < 2^m} (x + a) / prod_{a < 2^m} (2^m + a) at
    x = j + 1 - 2^m, where m is the number of trailing one bits of j;
  * logWalsh is the Walsh-Hadamard transform of logLUT (logLUT[0] := 0)
    modulo the field's modulus;
  * Leopard parity: with m = ceilPow2(parityShards), data shard g*m + r is
    the value at the point (g+1)*m + r of a polynomial P_g of degree < m
    (zero where g*m + r >= k), and parity i is sum_g P_g(i).  That is Lagrange
    interpolation, which depends neither on the FFT nor on the basis of the
    polynomial space; checked on small shapes against the oracle's encoder
    (the GPU parity tests compare the engine with the oracle byte for byte).

A basis with one constant replaced by the other root of x^2 + x = b[i-1]
(for example 0xACCA -> 0xACCB) still satisfies the Cantor property and
still gives a field and an MDS code; the brute-force products catch it
(test_other_root_of_a_basis_constant_is_caught).
"""
import json
import os

import numpy as np
import pytest

import reedsolomon_amd as rs
from oracle import rsref


@pytest.fixture(scope="module")
def consts(golden_dir):
    with open(os.path.join(golden_dir, "leopard_consts.json")) as f:
        return json.load(f)


def _clmul_mod(a, b, poly, bits):
    """Carry-less a*b reduced modulo `poly` (degree `bits`), elementwise."""
    a = np.asarray(a, np.int64)
    b = np.asarray(b, np.int64)
    r = np.zeros(np.broadcast(a, b).shape, np.int64)
    for i in range(bits):
        r ^= np.where((b >> i) & 1, a << i, 0)
    for i in range(2 * bits - 2, bits - 1, -1):
        r = np.where((r >> i) & 1, r ^ (poly << (i - bits)), r)
    return r


def _powers_of_x(poly, bits):
    """x^i mod poly for i < 2^bits - 1 (multiplying by x is a shift and a
    conditional reduction)."""
    modulus = (1 << bits) - 1
    out = [0] * modulus
    state = 1
    for i in range(modulus):
        out[i] = state
        state <<= 1
        if state >> bits:
            state ^= poly
    assert state == 1  # x^modulus = 1
    return np.array(out, np.int64)


def _phi(basis):
    """phi[a] = XOR of basis[i] for the set bits of a (Cantor coordinates ->
    polynomial basis), for every a of the field."""
    bits = len(basis)
    phi = np.zeros(1 << bits, np.int64)
    for i, b in enumerate(basis):
        w = 1 << i
        phi[w:2 * w] = phi[:w] ^ b
    return phi


def _inverse_perm(phi):
    inv = np.full(len(phi), -1, np.int64)
    inv[phi] = np.arange(len(phi))
    assert (inv >= 0).all(), "basis is not linearly independent"
    return inv


def _field(consts, bits):
    c = consts[f"gf{bits}"]
    return c["polynomial"], c["cantor_basis"], c["modulus"]


def _engine_tables(bits):
    return rs.leopard_tables(bits)


def _oracle_tables(bits):
    return rsref.leo16_tables() if bits == 16 else rsref.leo8_tables()


def test_fixture_constants(consts):
    assert consts["gf16"]["bitwidth"] == 16 and consts["gf16"]["polynomial"] == 0x1002D
    assert consts["gf16"]["cantor_basis"][:2] == [0x0001, 0xACCA]
    assert consts["gf8"]["bitwidth"] == 8 and consts["gf8"]["polynomial"] == 0x11D
    assert consts["gf8"]["cantor_basis"] == [1, 214, 152, 146, 86, 200, 88, 230]


@pytest.mark.parametrize("bits", [16, 8])
def test_cantor_basis_property(consts, bits):
    # beta_0 = 1 and beta_i^2 + beta_i = beta_{i-1} in GF(2)[x]/polynomial.
    poly, basis, modulus = _field(consts, bits)
    assert basis[0] == 1
    for i in range(1, bits):
        b = basis[i]
        assert int(_clmul_mod(b, b, poly, bits)) ^ b == basis[i - 1], i
    _inverse_perm(_phi(basis))  # a basis of the whole field
    # x generates the multiplicative group (the LFSR of leopard.go:1116-1123
    # visits every nonzero element once).
    pw = _powers_of_x(poly, bits)
    assert len(np.unique(pw)) == modulus and (pw > 0).all()
    assert (_clmul_mod(pw[:-1], 2, poly, bits) == pw[1:]).all()


@pytest.mark.parametrize("who", ["oracle", "engine"])
@pytest.mark.parametrize("bits", [16, 8])
def test_exp_log_tables_are_powers_of_x_in_cantor_coordinates(consts, bits, who):
    poly, basis, modulus = _field(consts, bits)
    t = _oracle_tables(bits) if who == "oracle" else _engine_tables(bits)
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    inv_phi = _inverse_perm(_phi(basis))
    pw = _powers_of_x(poly, bits)
    assert (exp[:modulus] == inv_phi[pw]).all()
    assert exp[modulus] == exp[0]
    assert log[0] == modulus
    assert (log[exp[:modulus]] == np.arange(modulus)).all()
    # Each basis constant is the polynomial form of the Cantor coordinate
    # vector 1 << i, read back from the tables.
    for i in range(bits):
        assert pw[log[1 << i]] == basis[i], i


def _brute_products(consts, bits, a, b):
    poly, basis, _ = _field(consts, bits)
    phi = _phi(basis)
    return _inverse_perm(phi)[_clmul_mod(phi[a], phi[b], poly, bits)]


def _pairs(bits, n, seed):
    rng = np.random.default_rng(seed)
    a = rng.integers(0, 1 << bits, n)
    b = rng.integers(0, 1 << bits, n)
    e = np.array([0] + [1 << i for i in range(bits)] + [(1 << bits) - 1])
    ea, eb = np.meshgrid(e, e)
    return np.concatenate([a, ea.ravel()]), np.concatenate([b, eb.ravel()])


def test_gf16_mul_brute_force(consts):
    a, b = _pairs(16, 100_000, 11)
    want = _brute_products(consts, 16, a, b)
    # Oracle: the table product of gf16.go:6-16 on every pair, the C entry
    # point on a sample.
    t = rsref.leo16_tables()
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    s = (log[a] + log[b]) % 65535
    got = np.where((a == 0) | (b == 0), 0, exp[s])
    assert (got == want).all()
    for i in range(0, len(a), 50):
        assert rsref.gf16_mul(int(a[i]), int(b[i])) == want[i]
    # Engine: LowLevel.GF16Mul on every pair.
    mul = rs.lib.rsx_gf16_mul
    bad = [(int(x), int(y)) for x, y, w in zip(a, b, want) if mul(int(x), int(y)) != w]
    assert not bad, bad[:5]


@pytest.mark.parametrize("who", ["oracle", "engine"])
def test_gf8_leopard_mul_brute_force(consts, who):
    # Every product of the 8-bit field (mulLog8(a, log b) = a*b).
    t = _oracle_tables(8) if who == "oracle" else _engine_tables(8)
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    a, b = np.meshgrid(np.arange(256), np.arange(256))
    a, b = a.ravel(), b.ravel()
    want = _brute_products(consts, 8, a, b)
    got = np.where((a == 0) | (b == 0), 0, exp[(log[a] + log[b]) % 255])
    assert (got == want).all()
    if who == "oracle":
        for x, y in zip(a[::97], b[::97]):
            got1 = rsref.leo8_mul_log(int(x), int(log[y])) if y else 0
            assert got1 == want[x + 256 * y]


def test_other_root_of_a_basis_constant_is_caught(consts):
    # beta and beta + 1 both solve x^2 + x = beta_{i-1}: swapping one keeps
    # the Cantor property, and the brute-force products then disagree with
    # the engine's on most pairs.
    poly, basis, _ = _field(consts, 16)
    wrong = list(basis)
    wrong[1] ^= 1  # 0xACCA -> 0xACCB
    assert int(_clmul_mod(wrong[1], wrong[1], poly, 16)) ^ wrong[1] == wrong[0]
    phi = _phi(wrong)
    a, b = _pairs(16, 2000, 5)
    want_wrong = _inverse_perm(phi)[_clmul_mod(phi[a], phi[b], poly, 16)]
    got = np.array([rs.lib.rsx_gf16_mul(int(x), int(y)) for x, y in zip(a, b)])
    assert (got != want_wrong).mean() > 0.5


def _subspace_skews(exp, log, bits):
    """log of W_m(j + 1 - 2^m) for every j < modulus (m = trailing ones of j)."""
    modulus = (1 << bits) - 1

    def mul(x, y):
        return np.where((x == 0) | (y == 0), 0, exp[(log[x] + log[y]) % modulus])

    out = np.zeros(modulus, np.int64)
    for m in range(bits):
        j = np.arange((1 << m) - 1, modulus, 1 << (m + 1))
        x = j + 1 - (1 << m)
        num = np.ones_like(x)
        den = np.ones(1, np.int64)
        for a in range(1 << m):
            num = mul(num, x ^ a)
            den = mul(den, np.array([(1 << m) ^ a]))
        inv_den = exp[(modulus - log[den]) % modulus]
        w = mul(num, inv_den)
        out[j] = np.where(w == 0, modulus, log[w])
    return out


@pytest.mark.parametrize("who", ["oracle", "engine"])
@pytest.mark.parametrize("bits", [16, 8])
def test_fft_skew_is_the_subspace_polynomial(bits, who):
    # exp/log are pinned above; fftSkew follows from them and the basis.
    t = _oracle_tables(bits) if who == "oracle" else _engine_tables(bits)
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    want = _subspace_skews(exp, log, bits)
    assert (t["skew"].astype(np.int64) == want).all()


@pytest.mark.parametrize("who", ["oracle", "engine"])
@pytest.mark.parametrize("bits", [16, 8])
def test_log_walsh_is_the_hadamard_transform(bits, who):
    t = _oracle_tables(bits) if who == "oracle" else _engine_tables(bits)
    order, modulus = 1 << bits, (1 << bits) - 1
    L = t["log"].astype(np.int64).copy()
    L[0] = 0
    j = np.arange(order)
    rows = range(order) if bits == 8 else np.random.default_rng(3).integers(0, order, 48)
    for i in rows:
        par = np.zeros(order, np.int64)
        x = j & int(i)
        while x.any():
            par ^= x & 1
            x >>= 1
        w = int((np.where(par == 1, -L, L).sum()) % modulus)
        assert int(t["walsh"][i]) % modulus == w, i


def _ceil_pow2(x):
    p = 1
    while p < x:
        p <<= 1
    return p


def _lagrange_parity(exp, log, bits, data_cols, k, p):
    """Parity symbols of one column by interpolation (see the module doc)."""
    modulus = (1 << bits) - 1

    def mul(x, y):
        return 0 if x == 0 or y == 0 else int(exp[(log[x] + log[y]) % modulus])

    def inv(x):
        return int(exp[(modulus - log[x]) % modulus])

    m = _ceil_pow2(p)
    out = [0] * p
    for g in range((k + m - 1) // m):
        pts = [(g + 1) * m + r for r in range(m)]
        for r in range(m):
            v = int(data_cols[g * m + r]) if g * m + r < k else 0
            if v == 0:
                continue
            den = 1
            for s in range(m):
                if s != r:
                    den = mul(den, pts[r] ^ pts[s])
            scale = mul(v, inv(den))
            for i in range(p):
                num = 1
                for s in range(m):
                    if s != r:
                        num = mul(num, i ^ pts[s])
                out[i] ^= mul(scale, num)
    return out


def _symbols(shard, bits):
    if bits == 8:
        return shard.astype(np.int64)
    c = shard.reshape(-1, 2, 32)  # leopard.go:907-925 split layout
    return (c[:, 0, :].astype(np.int64) | (c[:, 1, :].astype(np.int64) << 8)).reshape(-1)


SHAPES = [(1, 1), (2, 1), (3, 2), (4, 4), (5, 3), (10, 3), (8, 8), (17, 3), (3, 9), (33, 7),
          (20, 16)]


@pytest.mark.parametrize("bits", [16, 8])
@pytest.mark.parametrize("k,p", SHAPES)
def test_leopard_parity_is_lagrange_interpolation(bits, k, p):
    t = _oracle_tables(bits)
    exp, log = t["exp"].astype(np.int64), t["log"].astype(np.int64)
    rng = np.random.default_rng(bits * 1000 + k * 31 + p)
    size = 64
    shards = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + \
             [np.zeros(size, np.uint8) for _ in range(p)]
    assert rsref.Leopard(k, p, bits).encode(shards) == 0
    syms = [_symbols(s, bits) for s in shards]
    for col in range(0, len(syms[0]), 5):
        want = _lagrange_parity(exp, log, bits, [s[col] for s in syms[:k]], k, p)
        got = [int(syms[k + i][col]) for i in range(p)]
        assert got == want, (col, got, want)
