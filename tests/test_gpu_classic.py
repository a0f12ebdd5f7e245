"""GPU parity of the classic GF(2^8) engine against the CPU oracle.

Every parity and reconstructed byte must equal the oracle's (bit-exact).
Covers the reference's kernel cross-check shape sweep (galois_test.go:366-402,
every K x M <= 10 x 10 over several sizes), TestOneEncode, the matrix options,
EncodeIdx, Verify, all Reconstruct modes including every ReconstructSome
mask, error outcomes, unaligned device buffers, host and device memory, and
the RS(10,3) 1 MiB / RS(16,4) 4 MiB workloads.
"""
import json
import os

import numpy as np
import pytest

from oracle import rsref

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
import reedsolomon_amd as rs  # noqa: E402

SEED = 0xABADC0CAC01A  # reedsolomon_test.go:309, used as a seed label


def dev(a: np.ndarray):
    return torch.from_numpy(a.copy()).to("cuda")


def host(t) -> np.ndarray:
    return t.cpu().numpy()


def rand_data(rng, k, size):
    return [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]


def oracle_encode(k, m, data, kind=rsref.MATRIX_DEFAULT, custom=None):
    o = rsref.Classic(k, m, kind, custom)
    shards = [d.copy() for d in data] + [np.zeros(len(data[0]), np.uint8) for _ in range(m)]
    assert o.encode(shards) == 0
    return o, shards


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available() or rs.device_count() == 0:
        pytest.skip("no GPU")


def test_one_encode_kat(golden_dir):
    v = json.load(open(os.path.join(golden_dir, "kat.json")))["one_encode_5_5"]
    enc = rs.New(5, 5)
    for mode in ("host", "device"):
        shards = [np.array(d, np.uint8) for d in v["data"]] + [np.zeros(2, np.uint8) for _ in range(5)]
        if mode == "device":
            shards = [dev(s) for s in shards]
        enc.Encode(shards)
        got = [(host(s) if mode == "device" else s).tolist() for s in shards[5:]]
        assert got == v["parity"], mode
        assert enc.Verify(shards)


@pytest.mark.parametrize("size", [1, 31, 64, 1000, 4097])
def test_every_kxm_up_to_10x10(size):
    rng = np.random.default_rng(size)
    for k in range(1, 11):
        for m in range(1, 11):
            data = rand_data(rng, k, size)
            _, want = oracle_encode(k, m, data)
            enc = rs.New(k, m)
            shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda")
                                               for _ in range(m)]
            enc.Encode(shards)
            for i in range(m):
                assert np.array_equal(host(shards[k + i]), want[k + i]), (k, m, size, i)


# (20, 20), (200, 56), (1, 255): 16-output passes; (230, 26): too many
# inputs for a 16-output pass's tables, so 8-output passes.
@pytest.mark.parametrize("k,m", [(17, 3), (32, 8), (20, 20), (200, 56), (1, 255), (255, 1), (230, 26)])
def test_large_shapes_and_multi_pass(k, m):
    rng = np.random.default_rng(k * 1000 + m)
    size = 4096 + 48
    data = rand_data(rng, k, size)
    o, want = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
    enc.Encode(shards)
    for i in range(m):
        assert np.array_equal(host(shards[k + i]), want[k + i]), i
    lost = sorted(rng.choice(k + m, min(m, 24), replace=False).tolist())  # 9+ outputs: wide passes
    work = [None if i in lost else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert np.array_equal(host(work[i]), want[i]), i


@pytest.mark.parametrize("opt,kind", [(rs.WithCauchyMatrix, rsref.MATRIX_CAUCHY),
                                      (rs.WithPAR1Matrix, rsref.MATRIX_PAR1),
                                      (rs.WithJerasureMatrix, rsref.MATRIX_JERASURE)])
def test_matrix_options(opt, kind):
    rng = np.random.default_rng(kind)
    k, m, size = 8, 4, 777
    data = rand_data(rng, k, size)
    _, want = oracle_encode(k, m, data, kind)
    enc = rs.New(k, m, opt())
    assert np.array_equal(enc.CodingMatrix(), rsref.codec_matrix(k, m, kind))
    shards = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
    enc.Encode(shards)
    for i in range(m):
        assert np.array_equal(shards[k + i], want[k + i])


def test_custom_and_fast_one_parity_matrix():
    rng = np.random.default_rng(12)
    k, m, size = 6, 2, 300
    custom = rng.integers(1, 256, (m, k), dtype=np.uint8)
    data = rand_data(rng, k, size)
    _, want = oracle_encode(k, m, data, custom=custom)
    enc = rs.New(k, m, rs.WithCustomMatrix(custom))
    shards = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
    enc.Encode(shards)
    assert all(np.array_equal(shards[k + i], want[k + i]) for i in range(m))
    enc1 = rs.New(k, 1, rs.WithFastOneParityMatrix())
    s1 = [d.copy() for d in data] + [np.zeros(size, np.uint8)]
    enc1.Encode(s1)
    x = np.zeros(size, np.uint8)
    for d in data:
        x ^= d
    assert np.array_equal(s1[k], x)


def test_config_rs10_3_1mib_erase_3_and_7():
    # RS(10,3), 10 x 1 MiB random data, Encode + Verify, then data shards 3 and 7
    # set to nil and Reconstruct; parity and recovered bytes equal the oracle.
    rng = np.random.default_rng(SEED)
    k, m, size = 10, 3, 1 << 20
    data = rand_data(rng, k, size)
    o, want = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    for mode in ("device", "host"):
        shards = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
        if mode == "device":
            shards = [dev(s) for s in shards]
        enc.Encode(shards)
        assert enc.Verify(shards)
        for i in range(m):
            got = host(shards[k + i]) if mode == "device" else shards[k + i]
            assert np.array_equal(got, want[k + i]), (mode, i)
        shards[3] = None
        shards[7] = None
        enc.Reconstruct(shards)
        for i in (3, 7):
            got = host(shards[i]) if mode == "device" else shards[i]
            assert np.array_equal(got, data[i]), (mode, i)


def test_reconstruct_modes_vs_oracle():
    rng = np.random.default_rng(21)
    k, m, size = 9, 5, 2000
    data = rand_data(rng, k, size)
    o, full = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    for trial in range(20):
        e = int(rng.integers(1, m + 1))
        lost = set(rng.choice(k + m, e, replace=False).tolist())
        for method in ("Reconstruct", "ReconstructData"):
            ref = [None if i in lost else full[i].copy() for i in range(k + m)]
            err = getattr(o, "reconstruct" if method == "Reconstruct" else "reconstruct_data")(ref)
            assert err == 0
            work = [None if i in lost else dev(full[i]) for i in range(k + m)]
            getattr(enc, method)(work)
            for i in range(k + m):
                if ref[i] is None:
                    assert work[i] is None
                else:
                    assert np.array_equal(host(work[i]), ref[i]), (method, i, lost)
    assert enc.DecodeCacheSize() > 0


def test_inversion_cache_off():
    # WithInversionCache(false) (options.go:152-157): every reconstruct
    # inverts afresh and nothing is cached; results equal the oracle's,
    # also when the same pattern repeats back to back (the previous call's
    # plan is released while this one runs).
    rng = np.random.default_rng(22)
    k, m, size = 6, 3, 4096
    data = rand_data(rng, k, size)
    o, full = oracle_encode(k, m, data)
    enc = rs.New(k, m, rs.WithInversionCache(False))
    patterns = [{0}, {0}, {1, 7}, {2, 4, 8}, {1, 7}, {5}]
    works = []
    for lost in patterns:
        ref = [None if i in lost else full[i].copy() for i in range(k + m)]
        assert o.reconstruct(ref) == 0
        work = [None if i in lost else dev(full[i]) for i in range(k + m)]
        enc.Reconstruct(work)
        works.append((lost, ref, work))
    for lost, ref, work in works:
        for i in range(k + m):
            assert np.array_equal(host(work[i]), ref[i]), (lost, i)
    assert enc.DecodeCacheSize() == 0


def test_reconstruct_some_every_mask():
    # TestRSReconstructSome (reedsolomon_test.go:2319-2424)
    rng = np.random.default_rng(31)
    k, m = 3, 2
    data = rand_data(rng, k, 100)
    o, full = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    # A k-entry mask with a parity shard missing ({0, 4}): the reference
    # panics there; here the parity counts as not required.
    for req_len, lost in ((k, {0, 2}), (k, {0, 4}), (k + m, {0, 4})):
        for mask in range(1 << req_len):
            req = [bool(mask >> i & 1) for i in range(req_len)]
            ref = [None if i in lost else full[i].copy() for i in range(k + m)]
            assert o.reconstruct_some(ref, req) == 0
            work = [None if i in lost else dev(full[i]) for i in range(k + m)]
            enc.ReconstructSome(work, req)
            for i in range(k + m):
                assert (work[i] is None) == (ref[i] is None), (mask, i)
                if ref[i] is not None:
                    assert np.array_equal(host(work[i]), ref[i])


def test_encode_idx_matches_encode():
    rng = np.random.default_rng(41)
    k, m, size = 7, 4, 5000
    data = rand_data(rng, k, size)
    _, want = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    for mode in ("device", "host"):
        parity = [np.zeros(size, np.uint8) for _ in range(m)]
        if mode == "device":
            parity = [dev(p) for p in parity]
        for idx in range(k):
            enc.EncodeIdx(dev(data[idx]) if mode == "device" else data[idx], idx, parity)
        for i in range(m):
            got = host(parity[i]) if mode == "device" else parity[i]
            assert np.array_equal(got, want[k + i]), (mode, i)


def test_unaligned_device_buffers():
    rng = np.random.default_rng(51)
    k, m, size = 5, 3, 1234
    data = rand_data(rng, k, size)
    _, want = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    pool = torch.zeros((k + m) * (size + 3) + 7, dtype=torch.uint8, device="cuda")
    shards = []
    for i in range(k + m):
        off = 3 + i * (size + 3)
        t = pool[off:off + size]
        if i < k:
            t.copy_(torch.from_numpy(data[i]))
        shards.append(t)
    enc.Encode(shards)
    for i in range(m):
        assert np.array_equal(host(shards[k + i]), want[k + i])
    assert enc.Verify(shards)


def test_verify_detects_corruption_and_errors():
    rng = np.random.default_rng(61)
    k, m, size = 6, 3, 4096
    data = rand_data(rng, k, size)
    _, full = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    shards = [dev(s) for s in full]
    assert enc.Verify(shards)
    shards[k + 1][17] ^= 1
    assert not enc.Verify(shards)
    with pytest.raises(rs.ErrTooFewShards):
        enc.Encode(shards[:-1])
    bad = [dev(s) for s in full]
    bad[2] = bad[2][:100]
    with pytest.raises(rs.ErrShardSize):
        enc.Encode(bad)
    with pytest.raises(rs.ErrShardNoData):
        enc.Encode([None] * (k + m))
    work = [None] * 4 + [dev(s) for s in full[4:]]
    with pytest.raises(rs.ErrTooFewShards):
        enc.Reconstruct(work)
    with pytest.raises(rs.ErrMixedMemory):
        enc.Encode([dev(s) for s in full[:k]] + [s.copy() for s in full[k:]])


# (k, m, size, byte offset of the shards in their buffer, memory): the fused
# Verify pass recomputes parity in registers and compares it with the stored
# parity; corrupt one byte in each region a pass covers and alternate good and
# corrupt calls, so a flag left over from one call cannot decide the next.
VERIFY_CASES = [
    (10, 4, 1 << 20, 0, "dev"),       # wave chunks only
    (10, 3, 4096 + 16 * 37, 0, "dev"),  # chunks + per-vector tail
    (230, 20, 4096, 0, "dev"),        # three 8-output passes
    (6, 3, 4096 + 21, 0, "dev"),      # + byte tail (len % 16)
    (5, 2, 1000, 3, "dev"),           # unaligned shards: byte path only
    (10, 12, 8192, 0, "dev"),         # one 16-output pass with 12 outputs
    (4, 2, 5000, 0, "host"),          # staged through HBM
]


@pytest.mark.parametrize("k,m,size,shift,mem", VERIFY_CASES)
def test_fused_verify_finds_every_corrupt_byte(k, m, size, shift, mem):
    rng = np.random.default_rng(size + k + m)
    data = rand_data(rng, k, size)
    _, full = oracle_encode(k, m, data)
    enc = rs.New(k, m)
    if mem == "host":
        shards = [s.copy() for s in full]
    else:
        pool = torch.zeros((k + m) * (size + shift) + shift, dtype=torch.uint8, device="cuda")
        shards = []
        for i in range(k + m):
            t = pool[shift + i * (size + shift):shift + i * (size + shift) + size]
            t.copy_(torch.from_numpy(full[i]))
            shards.append(t)
    assert enc.Verify(shards)
    spots = sorted({0, size - 1, size // 2, (size // 16) * 16 - 1 if size >= 16 else 0, size - size % 16})
    for p in range(m):
        for b in spots:
            if b >= size:
                continue
            shards[k + p][b] ^= 0x40
            assert not enc.Verify(shards), (p, b)
            shards[k + p][b] ^= 0x40
            assert enc.Verify(shards), (p, b)
    # a corrupt data byte fails too
    shards[k // 2][size // 3] ^= 1
    assert not enc.Verify(shards)
    shards[k // 2][size // 3] ^= 1
    assert enc.Verify(shards)


def test_rs16_4_4mib_stripes_round_trip():
    # The RS(16,4) 4 MiB workload on a few stripes: the first stripe bit-exact
    # against the oracle, every stripe through encode -> erase -> reconstruct.
    rng = np.random.default_rng(SEED + 1)
    k, m, size = 16, 4, 4 << 20
    enc = rs.New(k, m)
    for stripe in range(3):
        data = rand_data(rng, k, size)
        shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda")
                                           for _ in range(m)]
        enc.Encode(shards)
        if stripe == 0:
            _, want = oracle_encode(k, m, data)
            for i in range(m):
                assert np.array_equal(host(shards[k + i]), want[k + i])
        orig = [s.clone() for s in shards]
        e = int(rng.integers(1, m + 1))
        lost = rng.choice(k + m, e, replace=False).tolist()
        work = [None if i in lost else shards[i] for i in range(k + m)]
        enc.Reconstruct(work)
        for i in lost:
            assert torch.equal(work[i], orig[i]), (stripe, i)


@pytest.mark.parametrize("k,m", [(5, 3), (16, 4), (9, 7)])
def test_long_shards_stream_path(k, m):
    # Shards of 16 MiB and more take the four-chunks-per-wave streaming path;
    # a ragged length also exercises the per-vector and per-byte tails.
    rng = np.random.default_rng(k * 31 + m)
    size = (16 << 20) + 1040 + 3
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    rows = rsref.codec_matrix(k, m)[k:]
    want = [np.zeros(size, np.uint8) for _ in range(m)]
    rsref.code_some_shards_fast(rows, data, want, threads=8)
    enc = rs.New(k, m)
    shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
    enc.Encode(shards)
    for i in range(m):
        assert np.array_equal(host(shards[k + i]), want[i]), i
    lost = sorted(rng.choice(k + m, m, replace=False).tolist())
    orig = {i: shards[i].clone() for i in lost}
    work = [None if i in lost else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert torch.equal(work[i], orig[i]), i


def test_one_codec_shared_by_threads():
    # An Encoder is safe for concurrent use (reedsolomon_test.go:2158-2196):
    # several threads encode and reconstruct different stripes with one codec.
    import threading
    k, m, size = 10, 4, 1 << 16
    enc = rs.New(k, m)
    o = rsref.Classic(k, m)
    rng = np.random.default_rng(77)
    stripes = []
    for _ in range(8):
        data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
        ref = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
        o.encode(ref)
        stripes.append(ref)
    errors = []

    def work(t):
        try:
            for rep in range(5):
                ref = stripes[(t + rep) % len(stripes)]
                shards = [torch.from_numpy(x.copy()).to("cuda") for x in ref[:k]] + [
                    torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
                enc.Encode(shards)
                lost = [(t + rep) % (k + m), (t + rep + 3) % (k + m)]
                for i in lost:
                    shards[i] = None
                enc.Reconstruct(shards)
                for i in range(k + m):
                    if not np.array_equal(shards[i].cpu().numpy(), ref[i]):
                        errors.append((t, rep, i))
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:5]
