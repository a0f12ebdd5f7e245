"""GPU parity of the Leopard GF(2^16) and GF(2^8) engines against the oracle.

Parity bytes must equal the oracle's (a line-faithful restatement of
leopard.go / leopard8.go); reconstructed bytes must equal the oracle's and the
original data.  Leopard is column-separable per 64-byte chunk, so the
k=4096, m=1024, 64 KiB workload is checked against the oracle on a column
slice of every shard and by a full encode -> erase 1024 -> reconstruct
round trip.
"""
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


def host(t):
    return t.cpu().numpy()


def opt(bits):
    return rs.WithLeopardGF16(True) if bits == 16 else rs.WithLeopardGF(True)


def oracle_encode(bits, k, m, data):
    o = rsref.Leopard(k, m, bits)
    shards = [d.copy() for d in data] + [np.zeros(len(data[0]), np.uint8) for _ in range(m)]
    assert o.encode(shards) == 0
    return o, shards


SHAPES = [(16, 1, 1), (16, 2, 1), (16, 3, 2), (16, 4, 4), (16, 5, 3), (16, 10, 3), (16, 8, 8),
          (16, 17, 3), (16, 33, 7), (16, 3, 9), (16, 100, 20), (16, 260, 40), (16, 300, 300),
          (8, 1, 1), (8, 3, 2), (8, 4, 4), (8, 10, 3), (8, 17, 3), (8, 33, 7), (8, 64, 64),
          (8, 100, 20), (8, 192, 56), (8, 128, 128)]


@pytest.mark.parametrize("bits,k,m", SHAPES)
def test_encode_reconstruct_vs_oracle(bits, k, m):
    rng = np.random.default_rng(bits * 10000 + k * 10 + m)
    size = 64 * int(rng.integers(1, 5))
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(bits, k, m, data)
    enc = rs.New(k, m, opt(bits))
    assert enc.ShardSizeMultiple() == 64
    for mode in ("device", "host"):
        shards = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
        if mode == "device":
            shards = [dev(s) for s in shards]
        enc.Encode(shards)
        for i in range(m):
            got = host(shards[k + i]) if mode == "device" else shards[k + i]
            assert np.array_equal(got, want[k + i]), (mode, i)
        assert enc.Verify(shards)
    for trial in range(3):
        e = int(rng.integers(1, m + 1))
        lost = set(rng.choice(k + m, e, replace=False).tolist())
        for method, ometh in (("Reconstruct", "reconstruct"), ("ReconstructData", "reconstruct_data")):
            ref = [None if i in lost else want[i].copy() for i in range(k + m)]
            assert getattr(o, ometh)(ref) == 0
            work = [None if i in lost else dev(want[i]) for i in range(k + m)]
            getattr(enc, method)(work)
            for i in range(k + m):
                assert (work[i] is None) == (ref[i] is None), (method, i)
                if ref[i] is not None:
                    assert np.array_equal(host(work[i]), ref[i]), (method, i, sorted(lost))
                    assert np.array_equal(ref[i], want[i])


@pytest.mark.parametrize("bits", [16, 8])
def test_verify_false_and_errors(bits):
    rng = np.random.default_rng(bits)
    k, m, size = 12, 4, 256
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    _, full = oracle_encode(bits, k, m, data)
    enc = rs.New(k, m, opt(bits))
    shards = [dev(s) for s in full]
    assert enc.Verify(shards)
    shards[k + 2][77] ^= 0x10
    assert not enc.Verify(shards)
    with pytest.raises(rs.ErrInvalidShardSize):
        enc.Encode([dev(np.zeros(100, np.uint8)) for _ in range(k + m)])
    with pytest.raises(rs.ErrTooFewShards):
        enc.Encode(shards[:-1])
    work = [None] * (m + 1) + [dev(s) for s in full[m + 1:]]
    with pytest.raises(rs.ErrTooFewShards):
        enc.Reconstruct(work)
    with pytest.raises(rs.ErrNotSupported):
        enc.EncodeIdx(dev(data[0]), 0, [dev(np.zeros(size, np.uint8)) for _ in range(m)])


def test_reconstruct_some_maps_to_reference_modes():
    # leopard.go:378-383: all shards only when len(required) == total.
    rng = np.random.default_rng(5)
    k, m, size = 6, 3, 128
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    _, full = oracle_encode(16, k, m, data)
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    work = [None, dev(full[1])] + [dev(s) for s in full[2:k]] + [None, dev(full[k + 1]), dev(full[k + 2])]
    enc.ReconstructSome(work, [True] * k)
    assert work[k] is None and np.array_equal(host(work[0]), full[0])
    work = [None, dev(full[1])] + [dev(s) for s in full[2:k]] + [None, dev(full[k + 1]), dev(full[k + 2])]
    enc.ReconstructSome(work, [True] * (k + m))
    assert np.array_equal(host(work[k]), full[k])


def test_leopard16_4096_1024_64kib():
    # The Leopard GF(2^16) workload: k=4096, m=1024, 64 KiB shards; erase a
    # seeded random 1024 shards and reconstruct via the FFT path.
    k, m, size = 4096, 1024, 64 << 10
    g = torch.Generator(device="cuda")
    g.manual_seed(4096)
    buf = torch.empty((k + m, size), dtype=torch.uint8, device="cuda")
    buf[:k].random_(0, 256, generator=g)
    buf[k:].zero_()
    shards = [buf[i] for i in range(k + m)]
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    enc.Encode(shards)
    # oracle on a 128-byte column slice of every shard (column-separable)
    sl = slice(64 * 100, 64 * 102)
    cols = buf[:, sl].cpu().numpy()
    _, want = oracle_encode(16, k, m, [cols[i].copy() for i in range(k)])
    for i in range(m):
        assert np.array_equal(cols[k + i], want[k + i]), i
    orig = buf.clone()
    rng = np.random.default_rng(1024)
    lost = sorted(rng.choice(k + m, m, replace=False).tolist())
    work = [None if i in set(lost) else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert torch.equal(work[i], orig[i]), i


def test_leopard16_three_pass_column_chunks():
    # n = 8192 reconstructs run the three passes per column chunk of at most
    # 64 KiB (512 MiB work arrays): 2 x 64 KiB + 192 bytes is three chunks, the
    # last one three 64-byte slices wide.
    k, m, size = 4096, 1024, 2 * (64 << 10) + 192
    g = torch.Generator(device="cuda")
    g.manual_seed(8192)
    buf = torch.empty((k + m, size), dtype=torch.uint8, device="cuda")
    buf[:k].random_(0, 256, generator=g)
    buf[k:].zero_()
    shards = [buf[i] for i in range(k + m)]
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    enc.Encode(shards)
    sl = slice(size - 128, size)  # the ragged last chunk against the oracle
    cols = buf[:, sl].cpu().numpy()
    _, want = oracle_encode(16, k, m, [cols[i].copy() for i in range(k)])
    for i in range(m):
        assert np.array_equal(cols[k + i], want[k + i]), i
    orig = buf.clone()
    rng = np.random.default_rng(8192)
    lost = sorted(rng.choice(k + m, m, replace=False).tolist())
    gone = set(lost)
    work = [None if i in gone else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert torch.equal(work[i], orig[i]), i


def test_leopard16_three_pass_encode_column_chunks():
    # m = ceilPow2(10000) = 16384: the encoder's three passes run per column
    # chunk of 32 KiB (512 MiB work arrays), two groups each; 32 KiB + 128
    # bytes is two chunks, the second 128 bytes wide.
    k, m, size = 20000, 10000, (32 << 10) + 128
    g = torch.Generator(device="cuda")
    g.manual_seed(16384)
    buf = torch.empty((k + m, size), dtype=torch.uint8, device="cuda")
    buf[:k].random_(0, 256, generator=g)
    buf[k:].zero_()
    shards = [buf[i] for i in range(k + m)]
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    enc.Encode(shards)
    for sl in (slice(0, 128), slice(size - 128, size)):
        cols = buf[:, sl].cpu().numpy()
        _, want = oracle_encode(16, k, m, [cols[i].copy() for i in range(k)])
        for i in range(m):
            assert np.array_equal(cols[k + i], want[k + i]), (sl, i)
    assert enc.Verify(shards)
    orig = buf.clone()
    rng = np.random.default_rng(16384)
    lost = sorted(rng.choice(k + m, m, replace=False).tolist())
    gone = set(lost)
    work = [None if i in gone else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert torch.equal(work[i], orig[i]), i


def test_leopard8_192_56_64kib():
    # The Leopard GF(2^8) slice of the served stream at (192, 56) 64 KiB
    # ((200, 56) indexes past the reference's tables).
    k, m, size = 192, 56, 64 << 10
    rng = np.random.default_rng(192)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(8, k, m, data)
    enc = rs.New(k, m, rs.WithLeopardGF(True))
    shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
    enc.Encode(shards)
    for i in range(m):
        assert np.array_equal(host(shards[k + i]), want[k + i]), i
    lost = set(rng.choice(k + m, m, replace=False).tolist())
    work = [None if i in lost else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    for i in lost:
        assert np.array_equal(host(work[i]), want[i]), i


@pytest.mark.parametrize("k,m", [(32768, 32768), (64512, 1024)])
def test_largest_in_field_gf16_shapes(k, m):
    # k + ceilPow2(m) = 65536: every field position is a shard.  (32768, 32768)
    # encodes one launch per layer and reconstructs n = 65536 rows in three
    # passes; (64512, 1024) encodes 63 groups in one LDS tile.
    size = 128
    rng = np.random.default_rng(k + m)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(16, k, m, data)
    buf = torch.zeros((k + m, size), dtype=torch.uint8, device="cuda")
    buf[:k] = torch.from_numpy(np.stack(data)).to("cuda")
    shards = [buf[i] for i in range(k + m)]
    enc = rs.New(k, m, opt(16))
    enc.Encode(shards)
    got = buf.cpu().numpy()
    for i in range(m):
        assert np.array_equal(got[k + i], want[k + i]), i
    assert enc.Verify(shards)
    lost = sorted(rng.choice(k + m, m, replace=False).tolist())
    gone = set(lost)
    work = [None if i in gone else shards[i] for i in range(k + m)]
    enc.Reconstruct(work)
    rebuilt = np.stack([host(work[i]) for i in lost])
    assert np.array_equal(rebuilt, np.stack([want[i] for i in lost]))


# Sparse erasures: the final FFT skips butterfly groups that hold no erased
# row; every schedule (one tile GF(2^8) / GF(2^16), quarter tile n = 4096,
# three passes n = 8192) against the oracle, per call and batched.
SPARSE_CASES = [(8, 192, 56, 64 * 37), (16, 100, 20, 64 * 41), (16, 1000, 100, 64 * 3),
                (16, 2000, 50, 64 * 2), (16, 2100, 1100, 128)]


@pytest.mark.parametrize("bits,k,m,size", SPARSE_CASES)
def test_sparse_erasures_prune_fft_exactly(bits, k, m, size):
    rng = np.random.default_rng(size + 7 * k)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(bits, k, m, data)
    enc = rs.New(k, m, opt(bits))
    n = k + m
    sets = [[0], [k - 1], [k], [n - 1], sorted(rng.choice(n, 2, replace=False).tolist()),
            sorted(rng.choice(k, 3, replace=False).tolist())]
    items = []
    for lost in sets:
        for data_only in (False, True):
            shards = [None if i in lost else dev(want[i]) for i in range(n)]
            (enc.ReconstructData if data_only else enc.Reconstruct)(shards)
            for i in lost:
                if data_only and i >= k:
                    continue
                assert np.array_equal(host(shards[i]), want[i]), (lost, data_only, i)
        work = [dev(want[i]) if i not in lost else torch.zeros(size, dtype=torch.uint8, device="cuda")
                for i in range(n)]
        lens_zero = list(lost)
        items.append((work, lens_zero))
    # batched: one plan with every sparse pattern (shapes the batch kernel takes)
    b = rs.Batch(0)
    plan_items = [(enc, rs.OP_RECONSTRUCT, [None if i in lz else w[i] for i in range(n)], w) for w, lz in items]
    status = b.run([(e, op, sh) for e, op, sh, _ in plan_items])
    for (e, op, sh, w), st, (_, lz) in zip(plan_items, status, items):
        if st == rs.ErrNotSupported.code:
            continue
        assert st == 0
        for i in lz:
            assert np.array_equal(host(sh[i]), want[i]), (lz, i)


# Three-pass reconstructs (n = 8192) skip pass-1 row blocks past inputCount
# (all zero), and pass-3 row blocks without an erased row (pass 2 does not
# store their rows).  Patterns: the last 1000 data shards erased (inputCount
# = m + 2000, so blocks from 4096 stay zero), parity only (outputs in the
# first block), one data shard in the middle, and a spread of both; per call
# and batched, against the oracle.
def test_three_pass_skips_empty_row_blocks():
    k, m, size = 3000, 2000, 128
    rng = np.random.default_rng(3000)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(16, k, m, data)
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    n = k + m
    sets = [list(range(2000, 3000)), [k, k + 7, n - 1], [1500],
            sorted(rng.choice(n, 300, replace=False).tolist())]
    items = []
    for lost in sets:
        ref = [None if i in set(lost) else want[i].copy() for i in range(n)]
        assert o.reconstruct(ref) == 0
        shards = [None if i in set(lost) else dev(want[i]) for i in range(n)]
        enc.Reconstruct(shards)
        for i in lost:
            assert np.array_equal(host(shards[i]), want[i]), (len(lost), i)
        w = [dev(want[i]) if i not in set(lost) else torch.zeros(size, dtype=torch.uint8, device="cuda")
             for i in range(n)]
        items.append(([None if i in set(lost) else w[i] for i in range(n)], w, lost))
    status = rs.Batch(0).run([(enc, rs.OP_RECONSTRUCT, sh) for sh, _, _ in items])
    for (sh, w, lost), st in zip(items, status):
        assert st == 0
        for i in lost:
            assert np.array_equal(host(sh[i]), want[i]), (len(lost), i)


# Shapes x sizes that give several column slices and a narrower last slice,
# run on both schedules: rows of a column slice in LDS (the default: one
# tile where the rows fit, a quarter-chunk tile for GF(2^16) reconstructs
# with n = 4096, three tiled passes for n >= 8192; GF(2^16) encodes with
# m = ceilPow2(parity) of 4096 and several groups in a quarter-chunk tile from
# 4 KiB shards, m >= 8192 and every m >= 2048 below 4 KiB in three tiled
# passes) and one launch per butterfly layer everywhere (RSX_LEOPARD_LAYERED=1).
SCHEDULE_CASES = [(8, 192, 56, 64 * 37), (8, 17, 3, 64 * 129), (8, 128, 128, 64 * 5), (8, 3, 2, 64 * 1000),
                  (16, 100, 20, 64 * 41), (16, 300, 300, 64 * 3), (16, 33, 7, 64 * 300), (16, 5, 3, 64 * 777),
                  (16, 1000, 100, 64 * 600), (16, 1500, 600, 64 * 5), (16, 2000, 50, 64 * 2),
                  (16, 1000, 100, 64 * 37), (16, 1100, 1000, 64 * 19),
                  (16, 2100, 1100, 64 * 2), (16, 1100, 1100, 64 * 3), (16, 1000, 3000, 64 * 3),
                  (16, 1024, 1024, 64 * 21), (16, 5000, 3000, 64 * 3), (16, 3000, 5000, 64 * 2),
                  (16, 9000, 5000, 64), (16, 5000, 3000, 4096 + 64 * 3), (16, 3000, 5000, 4096 + 64 * 2),
                  (16, 9000, 5000, 4096), (16, 20000, 10000, 64), (16, 8000, 16000, 128)]


@pytest.mark.parametrize("bits,k,m,size", SCHEDULE_CASES)
def test_lds_and_layered_schedules_match_oracle(bits, k, m, size, monkeypatch):
    rng = np.random.default_rng(size + k)
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    o, want = oracle_encode(bits, k, m, data)
    lost = set(rng.choice(k + m, m, replace=False).tolist())
    ref = [None if i in lost else want[i].copy() for i in range(k + m)]
    assert o.reconstruct(ref) == 0
    for layered in (False, True):
        if layered:
            monkeypatch.setenv("RSX_LEOPARD_LAYERED", "1")
        else:
            monkeypatch.delenv("RSX_LEOPARD_LAYERED", raising=False)
        enc = rs.New(k, m, opt(bits))
        shards = [dev(d) for d in data] + [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
        enc.Encode(shards)
        for i in range(m):
            assert np.array_equal(host(shards[k + i]), want[k + i]), (layered, i)
        assert enc.Verify(shards)
        # one corrupt byte at a time (last parity row's last byte, first parity
        # byte, a data byte), each followed by a clean call
        for i, b in ((k + m - 1, size - 1), (k, 0), (k // 2, size // 2)):
            shards[i][b] ^= 0x5A
            assert not enc.Verify(shards), (layered, i, b)
            shards[i][b] ^= 0x5A
            assert enc.Verify(shards), (layered, i, b)
        work = [None if i in lost else shards[i] for i in range(k + m)]
        enc.Reconstruct(work)
        for i in range(k + m):
            assert np.array_equal(host(work[i]), ref[i]), (layered, i)


@pytest.mark.parametrize("bits,k,m", [(16, 100, 20), (8, 32, 16)])
def test_one_leopard_codec_shared_by_threads(bits, k, m):
    # An Encoder is safe for concurrent use (reedsolomon_test.go:2158-2196):
    # threads encode, verify and reconstruct different stripes (different
    # pointer tables and erasure patterns) with one Leopard codec.
    import threading
    size = 64 * 9
    enc = rs.New(k, m, opt(bits))
    rng = np.random.default_rng(k + bits)
    stripes = [oracle_encode(bits, k, m, [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)])[1]
               for _ in range(6)]
    errors = []

    def work(t):
        try:
            for rep in range(4):
                ref = stripes[(t + rep) % len(stripes)]
                shards = [dev(x) for x in ref[:k]] + [torch.zeros(size, dtype=torch.uint8, device="cuda")
                                                      for _ in range(m)]
                enc.Encode(shards)
                if not enc.Verify(shards):
                    errors.append((t, rep, "verify"))
                lost = {(t * 7 + rep) % (k + m), (t + 5 * rep + 1) % (k + m), (3 * t + rep + 2) % (k + m)}
                for i in lost:
                    shards[i] = None
                enc.Reconstruct(shards)
                for i in range(k + m):
                    if not np.array_equal(host(shards[i]), ref[i]):
                        errors.append((t, rep, i))
        except Exception as e:  # surfaced below
            errors.append(repr(e))

    threads = [threading.Thread(target=work, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors[:5]


@pytest.mark.parametrize("k,m,size", [(300, 100, 256), (2000, 50, 128), (4096, 1024, 64)])
def test_new_patterns_error_locators_on_device(k, m, size):
    # A GF(2^16) pattern first met by a per-call reconstruct gets its error
    # locators and reveal logs from the device transform; a batch plan of the
    # same pattern on another codec computes them on the host.  Both must
    # rebuild the oracle's codeword exactly, for patterns from one erasure to
    # the maximum (every data shard or every parity shard among them).
    rng = np.random.default_rng(k + m)
    n = k + m
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    _, ref = oracle_encode(16, k, m, data)
    pats = [[0], list(range(m)), list(range(k, k + m)), sorted(rng.choice(n, m, replace=False).tolist()),
            sorted(rng.choice(n, int(rng.integers(2, m)), replace=False).tolist())]
    per_call = rs.New(k, m, opt(16))
    batched = rs.New(k, m, opt(16))
    for lost in pats:
        shards = [None if i in lost else dev(ref[i]) for i in range(n)]
        per_call.Reconstruct(shards)
        for i in range(n):
            assert np.array_equal(host(shards[i]), ref[i]), (lost[:4], i)
        work = [None if i in lost else dev(ref[i]) for i in range(n)]
        b = rs.Batch(0)
        assert b.run([(batched, rs.OP_RECONSTRUCT, work)]) == [0]
        for i in range(n):
            assert np.array_equal(host(work[i]), ref[i]), ("batched", lost[:4], i)
    # the same patterns again: the cached device blobs serve them
    for lost in pats[:2]:
        shards = [None if i in lost else dev(ref[i]) for i in range(n)]
        per_call.Reconstruct(shards)
        assert all(np.array_equal(host(shards[i]), ref[i]) for i in range(n))
