"""Pins the CPU oracle's Leopard GF(2^16) / GF(2^8) restatement.

The reference holds no Leopard parity vectors (only round trips), so the
oracle is pinned by:
  * leopard8.go:1271-1336 gf2p811dMulMatricesLeo8, the 8x8 bit matrix of
    x -> mulLog8(x, log) for every log: it fixes mulLog8 on all 256x256
    inputs and with it expLUT8/logLUT8;
  * the addMod known answers (leopard_test.go:167-190, leopard8_test.go:10-33);
  * GF(2^16) field axioms through GF16Mul (gf16_test.go) and GF(2)-linearity
    of mulLog;
  * encode -> erase -> reconstruct round trips (the decoder's error-locator /
    formal-derivative path is structurally independent of the encoder's
    grouped IFFT, so a wrong skew index or dropped group fails it), MDS over
    random k-subsets, Verify true/false, and the reference's error outcomes.
The field representation, the skews and the encoder's parity bytes are
pinned against mathematics in tests/test_leopard_field_pin.py.
"""
import json
import os

import numpy as np
import pytest

from oracle import rsref


def _bswap64(v: int) -> int:
    return int.from_bytes(int(v).to_bytes(8, "little"), "big")


def test_leo8_mul_log_pinned_by_reference_gfni_table(golden_dir):
    # galois_amd64_test.go:70-83 recipe: bit (outBit*8+inBit) of the matrix is
    # bit outBit of mulLog8(1<<inBit, log), then the 8 bytes are swapped.
    ref = np.load(os.path.join(golden_dir, "leo8_gfni.npy"))
    for lg in range(256):
        mtx = 0
        for in_bit in range(8):
            r = rsref.leo8_mul_log(1 << in_bit, lg)
            for out_bit in range(8):
                if (r >> out_bit) & 1:
                    mtx |= 1 << (out_bit * 8 + in_bit)
        assert _bswap64(mtx) == int(ref[lg]), lg
    # mulLog8 is GF(2)-linear in x, so the bit matrices fix it on all inputs.
    for lg in (0, 1, 77, 254, 255):
        for x in range(256):
            want = 0
            for b in range(8):
                if x >> b & 1:
                    want ^= rsref.leo8_mul_log(1 << b, lg)
            assert rsref.leo8_mul_log(x, lg) == want


def test_add_mod_kat(golden_dir):
    with open(os.path.join(golden_dir, "kat.json")) as f:
        kat = json.load(f)
    for a, b, want in kat["leo16_add_mod"]["cases"]:
        assert rsref.leo16_add_mod(a, b) == want
    for a, b, want in kat["leo8_add_mod"]["cases"]:
        assert rsref.leo8_add_mod(a, b) == want


def test_leo16_tables_structure():
    t = rsref.leo16_tables()
    # exp/log are inverse bijections on the 65535 nonzero elements.
    nz = np.arange(1, 65536)
    assert (t["exp"][t["log"][nz]] == nz).all()
    assert len(np.unique(t["log"][nz])) == 65535
    assert t["exp"][65535] == t["exp"][0]
    # mulLog(x, modulus) == x (leopard.go:984-996 with partial reduction).
    for x in (1, 2, 3, 0x1234, 0xFFFF):
        assert rsref.leo16_mul_log(x, 65535) == x


def test_gf16_field_axioms():
    # gf16_test.go: identity, zero, commutativity, associativity, distributivity, inverses.
    rng = np.random.default_rng(2)
    mul = rsref.gf16_mul
    for _ in range(300):
        a, b, c = (int(v) for v in rng.integers(0, 65536, 3))
        assert mul(a, 1) == a and mul(a, 0) == 0
        assert mul(a, b) == mul(b, a)
        assert mul(mul(a, b), c) == mul(a, mul(b, c))
        assert mul(a, b ^ c) == mul(a, b) ^ mul(a, c)
    t = rsref.leo16_tables()
    for a in (1, 2, 0x8000, 0xABCD, 0xFFFF):
        inv = int(t["exp"][(65535 - int(t["log"][a])) % 65535])
        assert mul(a, inv) == 1


def test_leo16_mul_log_linear():
    rng = np.random.default_rng(4)
    for _ in range(200):
        x, y, lg = (int(v) for v in rng.integers(0, 65536, 3))
        assert rsref.leo16_mul_log(x ^ y, lg) == rsref.leo16_mul_log(x, lg) ^ rsref.leo16_mul_log(y, lg)


def _rand(rng, k, m, size):
    return [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + \
           [np.zeros(size, np.uint8) for _ in range(m)]


SHAPES16 = [(1, 1), (2, 1), (3, 2), (4, 4), (5, 3), (10, 3), (8, 8), (17, 3), (33, 7), (3, 9),
            (100, 20), (260, 40)]
SHAPES8 = [(1, 1), (3, 2), (4, 4), (10, 3), (17, 3), (33, 7), (64, 64), (100, 20), (192, 56)]


@pytest.mark.parametrize("bits,k,m", [(16, k, m) for k, m in SHAPES16] + [(8, k, m) for k, m in SHAPES8])
def test_leopard_round_trip(bits, k, m):
    rng = np.random.default_rng(bits * 1000 + k * 7 + m)
    codec = rsref.Leopard(k, m, bits)
    size = 128
    shards = _rand(rng, k, m, size)
    data = [s.copy() for s in shards[:k]]
    assert codec.encode(shards) == 0
    for i in range(k):
        assert (shards[i] == data[i]).all()
    assert codec.verify(shards) == (0, True)
    orig = [s.copy() for s in shards]
    for trial in range(3):
        e = int(rng.integers(1, m + 1))
        lost = set(rng.choice(k + m, e, replace=False).tolist())
        work = [None if i in lost else s.copy() for i, s in enumerate(orig)]
        assert codec.reconstruct(work) == 0
        for i in range(k + m):
            assert (work[i] == orig[i]).all(), (i, lost)
        work = [None if i in lost else s.copy() for i, s in enumerate(orig)]
        assert codec.reconstruct_data(work) == 0
        for i in range(k):
            assert (work[i] == orig[i]).all()
    flipped = [s.copy() for s in orig]
    flipped[k + m - 1][3] ^= 0x40
    assert codec.verify(flipped) == (0, False)


def test_leopard_errors():
    rng = np.random.default_rng(9)
    for bits in (8, 16):
        codec = rsref.Leopard(4, 2, bits)
        shards = _rand(rng, 4, 2, 100)  # not a multiple of 64
        assert codec.encode(shards) == rsref.ERR_INVALID_SHARD_SIZE
        shards = _rand(rng, 4, 2, 64)
        assert codec.encode(shards[:5]) == rsref.ERR_TOO_FEW_SHARDS
        assert codec.encode(shards) == 0
        work = [None, None, None] + [s.copy() for s in shards[3:]]
        assert codec.reconstruct(work) == rsref.ERR_TOO_FEW_SHARDS


@pytest.mark.parametrize("k,p", [(65000, 536), (1, 65535), (64513, 1024)])
def test_leo16_out_of_field_shape_reported(k, p):
    # newFF16 (leopard.go:58) accepts k + p <= 65536, but with k + ceilPow2(p)
    # > 65536 the encoder reads past fftSkew (a panic in the reference): the
    # oracle reports ErrMaxShardNum for every entry point instead.
    codec = rsref.Leopard(k, p, 16)
    shards = [np.zeros(64, np.uint8) for _ in range(k + p)]
    assert codec.encode(shards) == rsref.ERR_MAX_SHARD_NUM
    assert codec.verify(shards)[0] == rsref.ERR_MAX_SHARD_NUM
    work = [None] + shards[1:]
    assert codec.reconstruct(work) == rsref.ERR_MAX_SHARD_NUM


def test_leo16_largest_in_field_shape_round_trips():
    # k + ceilPow2(p) = 65536 exactly: every position of the field is a shard.
    k, p, size = 32768, 32768, 64
    rng = np.random.default_rng(65536)
    shards = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + [np.zeros(size, np.uint8)
                                                                              for _ in range(p)]
    codec = rsref.Leopard(k, p, 16)
    assert codec.encode(shards) == 0
    assert codec.verify(shards) == (0, True)
    lost = set(rng.choice(k + p, p, replace=False).tolist())
    work = [None if i in lost else shards[i] for i in range(k + p)]
    assert codec.reconstruct(work) == 0
    assert all(np.array_equal(work[i], shards[i]) for i in range(k + p))


def test_leo8_invalid_shape_reported():
    # k + ceilPow2(m) > 256 indexes past the reference's tables (a panic there).
    codec = rsref.Leopard(200, 56, 8)
    shards = _rand(np.random.default_rng(0), 200, 56, 64)
    assert codec.encode(shards) == rsref.ERR_MAX_SHARD_NUM
