"""Pins the CPU oracle's classic GF(2^8) path against the reference.

Checks, in order of strength:
  * the oracle's derived tables equal the reference's literal tables
    (galois.go:29-848, 948) byte for byte;
  * the reference's known-answer tests (galois_test.go, matrix_test.go,
    reedsolomon_test.go TestOneEncode / TestBuildMatrixJerasure);
  * closed forms and invariants: field axioms, systematic identity top,
    MDS on every k x k row subset for small shapes, linearity, and
    reconstruct == original bytes for valid codewords.
"""
import itertools
import json
import os

import numpy as np
import pytest

from oracle import rsref


@pytest.fixture(scope="module")
def kat(golden_dir):
    with open(os.path.join(golden_dir, "kat.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ref_tables(golden_dir):
    return dict(np.load(os.path.join(golden_dir, "galois_tables.npz")))


def test_tables_match_reference_literals(ref_tables):
    t = rsref.gf8_tables()
    for name in ("exp", "log", "inv", "mul", "low", "high", "gfni"):
        np.testing.assert_array_equal(t[name], ref_tables[name], err_msg=name)


def test_gal_multiply_and_exp_kat(kat):
    for a, b, want in kat["gal_multiply"]["cases"]:
        assert rsref.gal_mul(a, b) == want
    for a, n, want in kat["gal_exp"]["cases"]:
        assert rsref.gal_exp(a, n) == want


def test_gal_mul_slice_kat(kat):
    v = kat["gal_mul_slice"]
    inp = np.array(v["input"], np.uint8)
    out = np.zeros_like(inp)
    for step in v["steps"]:
        rsref.gal_mul_slice(step["c"], inp, out, step["xor"])
        assert out.tolist() == step["expect"], step["c"]


def test_field_axioms_exhaustive_pairs():
    # galois_test.go:19-125: associativity/commutativity/distributivity/inverse.
    t = rsref.gf8_tables()
    mul = t["mul"].astype(np.int64)
    a = np.arange(256)
    assert (mul == mul.T).all()
    assert (mul[1] == a).all() and (mul[0] == 0).all()
    for x in range(1, 256):
        assert mul[x, t["inv"][x]] == 1
    rng = np.random.default_rng(1)
    x, y, z = rng.integers(0, 256, (3, 4096))
    assert (mul[mul[x, y], z] == mul[x, mul[y, z]]).all()
    assert (mul[x, y ^ z] == (mul[x, y] ^ mul[x, z])).all()


def test_matrix_kat(kat):
    mm = kat["matrix_multiply"]
    got = rsref.matrix_multiply(np.array(mm["a"]), np.array(mm["b"]))
    assert got.tolist() == mm["expect"]
    for case in kat["matrix_inverse"]["cases"]:
        err, inv = rsref.matrix_invert(np.array(case["m"], np.uint8))
        if "error" in case:
            assert err == {"not_square": rsref.ERR_NOT_SQUARE, "singular": rsref.ERR_SINGULAR}[case["error"]]
        else:
            assert err == 0 and inv.tolist() == case["expect"]


def test_jerasure_matrix_kat(kat):
    m = rsref.codec_matrix(8, 4, rsref.MATRIX_JERASURE)
    assert (m[:8] == np.eye(8, dtype=np.uint8)).all()
    assert m[8:].tolist() == kat["jerasure_8_4"]["parity_rows"]


def test_one_encode_kat(kat):
    v = kat["one_encode_5_5"]
    codec = rsref.Classic(5, 5)
    shards = [np.array(d, np.uint8) for d in v["data"]] + [np.zeros(2, np.uint8) for _ in range(5)]
    assert codec.encode(shards) == 0
    assert [s.tolist() for s in shards[5:]] == v["parity"]
    assert codec.verify(shards) == (0, True)


@pytest.mark.parametrize("kind", [rsref.MATRIX_DEFAULT, rsref.MATRIX_CAUCHY, rsref.MATRIX_JERASURE])
def test_mds_every_square_subset(kind):
    # TestStandardMatrices / TestCauchyMatrices (reedsolomon_test.go:1880-2068) on small shapes.
    for k, m in [(2, 2), (3, 3), (4, 2), (5, 3)]:
        mat = rsref.codec_matrix(k, m, kind)
        assert (mat[:k] == np.eye(k, dtype=np.uint8)).all()
        for rows in itertools.combinations(range(k + m), k):
            err, _ = rsref.matrix_invert(mat[list(rows)])
            assert err == 0, (k, m, rows)


def test_default_matrix_is_vandermonde_times_inverse_top():
    # Closed form of reedsolomon.go:232-256: M = V * inv(V[:k]), V[r][c] = r^c.
    k, m = 6, 4
    v = np.array([[rsref.gal_exp(r, c) for c in range(k)] for r in range(k + m)], np.uint8)
    err, inv = rsref.matrix_invert(v[:k])
    assert err == 0
    assert (rsref.matrix_multiply(v, inv) == rsref.codec_matrix(k, m)).all()


def test_par1_and_xor_matrices_closed_form():
    k = 5
    p1 = rsref.codec_matrix(k, 3, rsref.MATRIX_PAR1)
    for r in range(3):
        for c in range(k):
            assert p1[k + r, c] == rsref.gal_exp(c + 1, r)
    x = rsref.codec_matrix(k, 1, rsref.MATRIX_XOR)
    assert (x[k] == 1).all()


def _rand_shards(rng, k, m, size):
    return [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + \
           [np.zeros(size, np.uint8) for _ in range(m)]


def test_encode_is_linear_and_column_separable():
    rng = np.random.default_rng(7)
    codec = rsref.Classic(7, 3)
    a = _rand_shards(rng, 7, 3, 333)
    b = _rand_shards(rng, 7, 3, 333)
    c = [x ^ y for x, y in zip(a, b)]
    for s in (a, b, c):
        assert codec.encode(s) == 0
    for i in range(7, 10):
        assert (c[i] == (a[i] ^ b[i])).all()
    # parity of a byte column only depends on that column
    part = [s[100:200].copy() for s in a]
    assert codec.encode(part) == 0
    for i in range(7, 10):
        assert (part[i] == a[i][100:200]).all()


@pytest.mark.parametrize("k,m", [(10, 3), (4, 4), (17, 3), (1, 1)])
def test_reconstruct_recovers_original(k, m):
    rng = np.random.default_rng(k * 100 + m)
    codec = rsref.Classic(k, m)
    shards = _rand_shards(rng, k, m, 1000)
    assert codec.encode(shards) == 0
    orig = [s.copy() for s in shards]
    for trial in range(6):
        e = int(rng.integers(1, m + 1))
        lost = rng.choice(k + m, e, replace=False)
        work = [None if i in lost else s.copy() for i, s in enumerate(orig)]
        assert codec.reconstruct(work) == 0
        for i in range(k + m):
            assert (work[i] == orig[i]).all()
        work = [None if i in lost else s.copy() for i, s in enumerate(orig)]
        assert codec.reconstruct_data(work) == 0
        for i in range(k):
            assert (work[i] == orig[i]).all()
        for i in range(k, k + m):
            assert work[i] is None or i not in lost


def test_reconstruct_errors_and_verify_false():
    rng = np.random.default_rng(3)
    codec = rsref.Classic(5, 2)
    shards = _rand_shards(rng, 5, 2, 64)
    assert codec.encode(shards) == 0
    work = [None, None, None] + [s.copy() for s in shards[3:]]
    assert codec.reconstruct(work) == rsref.ERR_TOO_FEW_SHARDS
    assert codec.encode(shards[:6]) == rsref.ERR_TOO_FEW_SHARDS
    bad = [s.copy() for s in shards]
    bad[2] = bad[2][:10]
    assert codec.encode(bad) == rsref.ERR_SHARD_SIZE
    assert codec.encode([np.zeros(0, np.uint8)] * 7) == rsref.ERR_SHARD_NO_DATA
    flipped = [s.copy() for s in shards]
    flipped[6][5] ^= 1
    assert codec.verify(flipped) == (0, False)


def test_reconstruct_some_matches_required_mask():
    # TestRSReconstructSome (reedsolomon_test.go:2319-2424): every required mask.
    rng = np.random.default_rng(11)
    k, m = 3, 2
    codec = rsref.Classic(k, m)
    shards = _rand_shards(rng, k, m, 50)
    codec.encode(shards)
    for mask in range(1 << (k + m)):
        req = [bool(mask >> i & 1) for i in range(k + m)]
        work = [None, s1 := shards[1].copy(), None, shards[3].copy(), shards[4].copy()]
        del s1
        assert codec.reconstruct_some(work, req) == 0
        for i in (0, 2):
            if req[i]:
                assert (work[i] == shards[i]).all()
            else:
                assert work[i] is None


def test_reconstruct_some_data_mask_with_parity_missing():
    # A k-entry mask (ReconstructData of the required data shards) while a
    # parity shard is missing: only the required data shards are rebuilt.
    rng = np.random.default_rng(12)
    k, m = 3, 2
    codec = rsref.Classic(k, m)
    shards = _rand_shards(rng, k, m, 50)
    codec.encode(shards)
    for mask in range(1 << k):
        req = [bool(mask >> i & 1) for i in range(k)]
        work = [None, shards[1].copy(), shards[2].copy(), shards[3].copy(), None]
        assert codec.reconstruct_some(work, req) == 0
        assert work[4] is None
        if req[0]:
            assert (work[0] == shards[0]).all()
        else:
            assert work[0] is None


def test_encode_idx_equals_encode():
    rng = np.random.default_rng(5)
    k, m = 6, 3
    codec = rsref.Classic(k, m)
    shards = _rand_shards(rng, k, m, 200)
    codec.encode(shards)
    parity = [np.zeros(200, np.uint8) for _ in range(m)]
    for idx in range(k):
        assert codec.encode_idx(shards[idx], idx, parity) == 0
    for i in range(m):
        assert (parity[i] == shards[k + i]).all()


@pytest.mark.parametrize("isa", [0, 1, 2])
def test_fast_cpu_path_equals_scalar(isa):
    # The SIMD restatement used for the CPU baseline must give the scalar
    # oracle's bytes for every tile shape (k, m up to and beyond 10 x 10).
    if isa > rsref.fast_isa():
        pytest.skip("ISA not available on this host")
    rng = np.random.default_rng(isa)
    for k, m, size in [(1, 1, 31), (10, 4, 4097), (16, 4, 65536 + 77), (17, 3, 999), (32, 12, 2048)]:
        rows = rng.integers(0, 256, (m, k), dtype=np.uint8)
        data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
        want = [np.zeros(size, np.uint8) for _ in range(m)]
        rsref.code_some_shards(rows, data, want)
        got = [np.zeros(size, np.uint8) for _ in range(m)]
        rsref.code_some_shards_fast(rows, data, got, threads=3, isa=isa)
        for i in range(m):
            assert np.array_equal(got[i], want[i]), (k, m, size, i)
