"""GPU parity of the batched path: every item of a mixed plan must equal the
plain per-stripe call (and the oracle) byte for byte, and report the same
error the plain call reports."""
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


SHAPES = [(4, 2), (6, 3), (10, 4), (17, 3), (32, 8), (20, 12)]


def _stripe(rng, k, m, size, offset=0):
    pool = torch.zeros((k + m) * (size + offset) + 16, dtype=torch.uint8, device="cuda")
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    shards = []
    for i in range(k + m):
        t = pool[offset + i * (size + offset): offset + i * (size + offset) + size]
        if i < k:
            t.copy_(torch.from_numpy(data[i]))
        shards.append(t)
    return data, shards


def test_mixed_batch_matches_plain_and_oracle():
    rng = np.random.default_rng(2024)
    codecs = {s: rs.New(*s) for s in SHAPES}
    items, expect = [], []
    for n in range(60):
        k, m = SHAPES[n % len(SHAPES)]
        size = int(rng.choice([16, 64, 1000, 4096, 5000, 70000, 1 << 18]))
        offset = 1 if n % 7 == 0 else 0  # some unaligned stripes
        data, shards = _stripe(rng, k, m, size, offset)
        o = rsref.Classic(k, m)
        ref = [d.copy() for d in data] + [np.zeros(size, np.uint8) for _ in range(m)]
        assert o.encode(ref) == 0
        if n % 3 == 0:
            items.append((codecs[(k, m)], rs.OP_ENCODE, shards))
            expect.append(("enc", k, m, ref))
        else:
            # reconstruct from a valid codeword with random erasures
            for i in range(k, k + m):
                shards[i].copy_(torch.from_numpy(ref[i]))
            e = int(rng.integers(0, m + 1))
            lost = set(rng.choice(k + m, e, replace=False).tolist())
            work = [None if i in lost else shards[i] for i in range(k + m)]
            op = rs.OP_RECONSTRUCT if n % 3 == 1 else rs.OP_RECONSTRUCT_DATA
            items.append((codecs[(k, m)], op, work))
            expect.append(("rec" if op == rs.OP_RECONSTRUCT else "recdata", k, m, ref, lost))
    b = rs.Batch(0)
    status = b.run(items)
    assert all(s == 0 for s in status), status
    assert b.num_tasks() > 0
    for (enc, op, shards), exp in zip(items, expect):
        kind, k, m = exp[:3]
        ref = exp[3]
        if kind == "enc":
            for i in range(k, k + m):
                assert np.array_equal(shards[i].cpu().numpy(), ref[i]), ("enc", k, m, i)
        else:
            lost = exp[4]
            for i in range(k + m):
                if i in lost and (kind == "rec" or i < k):
                    assert shards[i] is not None, (kind, i)
                    assert np.array_equal(shards[i].cpu().numpy(), ref[i]), (kind, k, m, i)
                elif i in lost:
                    assert shards[i] is None
    # a plan may be replayed: outputs are overwritten with the same bytes
    b.launch()
    b.synchronize()
    for (enc, op, shards), exp in zip(items, expect):
        if exp[0] == "enc":
            k, m, ref = exp[1], exp[2], exp[3]
            for i in range(k, k + m):
                assert np.array_equal(shards[i].cpu().numpy(), ref[i])


def test_batch_item_errors_match_plain_calls():
    rng = np.random.default_rng(7)
    enc = rs.New(5, 2)
    _, good = _stripe(rng, 5, 2, 256)
    _, bad = _stripe(rng, 5, 2, 256)
    bad[3] = bad[3][:100]
    too_few = [None, None, None, good[3], good[4], good[5], good[6]]
    b = rs.Batch(0)
    status = b.run([(enc, rs.OP_ENCODE, good), (enc, rs.OP_ENCODE, bad),
                    (enc, rs.OP_RECONSTRUCT, too_few), (enc, rs.OP_ENCODE, good[:6])])
    assert status == [0, rs.ErrShardSize.code, rs.ErrTooFewShards.code, rs.ErrTooFewShards.code]


@pytest.mark.parametrize("tile,bpc", [(4096, 1), (65536, 4), (1 << 20, 8)])
def test_manual_tuning_gives_same_bytes(tile, bpc):
    rng = np.random.default_rng(tile)
    enc = rs.New(10, 4)
    data, shards = _stripe(rng, 10, 4, 300000)
    o = rsref.Classic(10, 4)
    ref = [d.copy() for d in data] + [np.zeros(300000, np.uint8) for _ in range(4)]
    o.encode(ref)
    b = rs.Batch(0)
    b.set_tuning(tile, bpc)
    assert b.run([(enc, rs.OP_ENCODE, shards)]) == [0]
    assert b.tuning() == (tile, bpc)
    for i in range(10, 14):
        assert np.array_equal(shards[i].cpu().numpy(), ref[i])


def test_batched_classic_verify_compare_mode():
    # Verify items run as compare-mode passes (parity recomputed in registers
    # and checked against the item's parity shards): good and corrupt items of
    # several shapes, byte tails, unaligned shards and more than 8 parity
    # shards (two passes) in one plan, replayed with and without the graph.
    rng = np.random.default_rng(919)
    shapes = [(10, 4, 1 << 16), (6, 3, 4096 + 21), (5, 2, 1000), (10, 12, 8192), (17, 3, 64 * 33)]
    items, want = [], []
    for k, m, size in shapes:
        enc = rs.New(k, m)
        data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
        full = data + [np.zeros(size, np.uint8) for _ in range(m)]
        rsref.Classic(k, m).encode(full)
        for corrupt in (None, "parity", "data", "tail"):
            shift = 3 if size == 1000 else 0  # unaligned shards
            pool = torch.zeros((k + m) * (size + shift) + shift, dtype=torch.uint8, device="cuda")
            sh = []
            for i in range(k + m):
                t = pool[shift + i * (size + shift):shift + i * (size + shift) + size]
                t.copy_(torch.from_numpy(full[i]))
                sh.append(t)
            if corrupt == "parity":
                sh[k + m - 1][size // 2] ^= 0x11
            elif corrupt == "data":
                sh[0][7] ^= 0x80
            elif corrupt == "tail":
                sh[k][size - 1] ^= 0x01
            items.append((enc, rs.OP_VERIFY, sh))
            want.append(corrupt is None)
    # an erroneous verify item (a missing parity shard) beside them, reported
    # as the plain call reports it
    bad_enc = rs.New(4, 2)
    bad_item = [torch.zeros(64, dtype=torch.uint8, device="cuda") for _ in range(5)] + [None]
    with pytest.raises(rs.ErrShardSize):
        bad_enc.Verify(bad_item)
    items.append((bad_enc, rs.OP_VERIFY, bad_item))
    b = rs.Batch(0)
    for use_graph in (True, False):
        b.set_graph(use_graph)
        for _ in range(2):  # a replay starts from cleared flags
            status = b.run(items)
            assert status[:-1] == [0] * (len(items) - 1), status
            assert status[-1] == rs.ErrShardSize.code
            for idx, ok in enumerate(want):
                assert b.verify_result(idx) == ok, idx


def _leo_codeword(rng, bits, k, m, size):
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    full = data + [np.zeros(size, np.uint8) for _ in range(m)]
    assert rsref.Leopard(k, m, bits).encode(full) == 0
    return full


# GF(2^16) (2000, 50) and (1100, 1000) reconstruct n = 4096 work rows in
# quarter-chunk tiles (a launch of their own inside the plan).
LEO_CASES = [(8, 192, 56, 65536), (8, 10, 3, 64 * 37), (8, 64, 64, 4096), (16, 100, 20, 64 * 41),
             (16, 300, 300, 192), (16, 33, 7, 64 * 300), (16, 1000, 100, 128), (16, 2000, 50, 64 * 5),
             (16, 1100, 1000, 64 * 3)]


def test_batched_leopard_items_match_oracle_and_plain():
    rng = np.random.default_rng(808)
    items, checks = [], []
    for bits, k, m, size in LEO_CASES:
        enc = rs.New(k, m, rs.WithLeopardGF16(True) if bits == 16 else rs.WithLeopardGF(True))
        full = _leo_codeword(rng, bits, k, m, size)
        dev = [torch.from_numpy(x).to("cuda") for x in full]
        # encode into stale parity
        shards = dev[:k] + [torch.from_numpy(rng.integers(0, 256, size, dtype=np.uint8)).to("cuda") for _ in range(m)]
        items.append((enc, rs.OP_ENCODE, shards))
        checks.append(("enc", k, m, full, shards))
        # verify: good and corrupted
        good = [x.clone() for x in dev]
        bad = [x.clone() for x in dev]
        bad[int(rng.integers(0, k + m))][int(rng.integers(0, size))] ^= 0x21
        items.append((enc, rs.OP_VERIFY, good))
        checks.append(("ver", True))
        items.append((enc, rs.OP_VERIFY, bad))
        checks.append(("ver", False))
        # reconstruct / reconstruct data
        for op in (rs.OP_RECONSTRUCT, rs.OP_RECONSTRUCT_DATA):
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            work = [None if i in lost else dev[i].clone() for i in range(k + m)]
            items.append((enc, op, work))
            checks.append(("rec", k, m, full, work, lost, op))
    # a classic item in the same plan
    cenc = rs.New(10, 4)
    cdata = [rng.integers(0, 256, 5000, dtype=np.uint8) for _ in range(10)]
    cref = cdata + [np.zeros(5000, np.uint8) for _ in range(4)]
    rsref.Classic(10, 4).encode(cref)
    cshards = [torch.from_numpy(x).to("cuda") for x in cdata] + [torch.zeros(5000, dtype=torch.uint8, device="cuda")
                                                               for _ in range(4)]
    items.append((cenc, rs.OP_ENCODE, cshards))
    checks.append(("enc", 10, 4, cref, cshards))
    b = rs.Batch(0)
    for use_graph in (True, False):
        b.set_graph(use_graph)
        status = b.run(items)
        assert status == [0] * len(items), status
        for idx, c in enumerate(checks):
            if c[0] == "enc":
                _, k, m, full, shards = c
                for i in range(k, k + m):
                    assert np.array_equal(shards[i].cpu().numpy(), full[i]), (idx, i)
            elif c[0] == "ver":
                assert b.verify_result(idx) == c[1], idx
            else:
                _, k, m, full, work, lost, op = c
                for i in range(k + m):
                    if i in lost and (op == rs.OP_RECONSTRUCT or i < k):
                        assert np.array_equal(work[i].cpu().numpy(), full[i]), (idx, i)
                    elif i in lost:
                        assert work[i] is None
    assert b.num_tasks() > 0


def test_batched_many_gf8_leopard_items_match_oracle():
    # Many GF(2^8) Leopard items per plan, so workgroups walk slices of many
    # tasks in turn (forward task search), encodes of several groups take the
    # pipelined path (groups' XOR in a second LDS region, the next group's rows
    # loading during the IFFT), and reconstructs, Verify and encodes of
    # different shapes share one launch.
    rng = np.random.default_rng(2025)
    shapes = [(192, 56, 64 * 40), (10, 3, 64 * 37), (100, 28, 64 * 9), (40, 24, 1024)]
    codecs = {(k, m): rs.New(k, m, rs.WithLeopardGF(True)) for k, m, _ in shapes}
    items, checks = [], []
    for n in range(24):
        k, m, size = shapes[n % len(shapes)]
        enc = codecs[(k, m)]
        full = _leo_codeword(rng, 8, k, m, size)
        dev = [torch.from_numpy(x).to("cuda") for x in full]
        kind = n % 3
        if kind == 0:  # encode into stale parity
            shards = dev[:k] + [torch.from_numpy(rng.integers(0, 256, size, dtype=np.uint8)).to("cuda")
                                for _ in range(m)]
            items.append((enc, rs.OP_ENCODE, shards))
            checks.append(("enc", k, m, full, shards))
        elif kind == 1:  # Verify, every other one corrupted
            good = n % 2 == 0
            shards = [x.clone() for x in dev]
            if not good:
                shards[int(rng.integers(0, k + m))][int(rng.integers(0, size))] ^= 0x5A
            items.append((enc, rs.OP_VERIFY, shards))
            checks.append(("ver", good))
        else:  # reconstruct 1..m erasures
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            work = [None if i in lost else dev[i].clone() for i in range(k + m)]
            items.append((enc, rs.OP_RECONSTRUCT, work))
            checks.append(("rec", k, m, full, work, lost))
    b = rs.Batch(0)
    for use_graph in (True, False):
        b.set_graph(use_graph)
        status = b.run(items)
        assert status == [0] * len(items), status
        for idx, c in enumerate(checks):
            if c[0] == "enc":
                _, k, m, full, shards = c
                for i in range(k, k + m):
                    assert np.array_equal(shards[i].cpu().numpy(), full[i]), (idx, i)
            elif c[0] == "ver":
                assert b.verify_result(idx) == c[1], idx
            else:
                _, k, m, full, work, lost = c
                for i in lost:
                    assert np.array_equal(work[i].cpu().numpy(), full[i]), (idx, i)


@pytest.mark.parametrize("k,m,size", [(5000, 3000, 64 * 2), (3000, 5000, 64 * 3)])
def test_batched_tall_gf16_encode_and_verify_match_oracle(k, m, size):
    # m = ceilPow2(parity) of 4096 with two groups, and 8192 with one: the
    # quarter-chunk launch of the plan.
    rng = np.random.default_rng(k)
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    full = _leo_codeword(rng, 16, k, m, size)
    dev = [torch.from_numpy(x).to("cuda") for x in full]
    stale = dev[:k] + [torch.zeros(size, dtype=torch.uint8, device="cuda") for _ in range(m)]
    good = [x.clone() for x in dev]
    bad = [x.clone() for x in dev]
    bad[k + m - 1][size - 1] ^= 0x40
    b = rs.Batch(0)
    st = b.run([(enc, rs.OP_ENCODE, stale), (enc, rs.OP_VERIFY, good), (enc, rs.OP_VERIFY, bad)])
    assert st == [0, 0, 0], st
    for i in range(k, k + m):
        assert np.array_equal(stale[i].cpu().numpy(), full[i]), i
    assert b.verify_result(1) and not b.verify_result(2)


def test_batched_three_pass_reconstructs_rebuild_the_codeword():
    # GF(2^16) reconstructs with 8192, 16384 and 32768 work rows (P = 512,
    # 1024, 512) in one plan beside a one-tile Leopard item and a classic
    # item: one launch per pass over all three-pass items.
    rng = np.random.default_rng(31)
    stripes = []
    for k, m, size, op in ((5000, 100, 128, rs.OP_RECONSTRUCT), (12000, 3000, 64, rs.OP_RECONSTRUCT_DATA),
                           (9000, 5000, 64, rs.OP_RECONSTRUCT), (100, 20, 64 * 3, rs.OP_RECONSTRUCT)):
        enc = rs.New(k, m, rs.WithLeopardGF16(True))
        full = _leo_codeword(rng, 16, k, m, size)
        stripes.append((enc, k, m, full, set(rng.choice(k + m, m, replace=False).tolist()), op))
    cdata = [rng.integers(0, 256, 4096, dtype=np.uint8) for _ in range(10)]
    cref = cdata + [np.zeros(4096, np.uint8) for _ in range(4)]
    rsref.Classic(10, 4).encode(cref)
    stripes.append((rs.New(10, 4), 10, 4, cref, {2, 11}, rs.OP_RECONSTRUCT))
    b = rs.Batch(0)
    for use_graph in (True, False):
        b.set_graph(use_graph)
        items = []
        for enc, k, m, full, lost, op in stripes:
            items.append((enc, op, [None if i in lost else torch.from_numpy(full[i]).to("cuda")
                                    for i in range(k + m)]))
        st = b.run(items)
        assert st == [0] * len(items), st
        for idx, ((enc, k, m, full, lost, op), (_, _, work)) in enumerate(zip(stripes, items)):
            for i in lost:
                if op == rs.OP_RECONSTRUCT or i < k:
                    assert np.array_equal(work[i].cpu().numpy(), full[i]), (use_graph, idx, i)
                else:
                    assert work[i] is None, (use_graph, idx, i)


def test_batched_three_pass_encodes_match_oracle():
    # GF(2^16) encodes above the tiles (m = 16384 with two groups and with
    # one, m = 8192 with two) run the encoder's passes per group index, in a
    # plan beside a one-tile encode and a three-pass reconstruct; Verify of
    # such a shape compares in pass 3.
    rng = np.random.default_rng(16)
    stripes = []
    for k, m, size in ((20000, 10000, 64), (8000, 16000, 128), (10000, 5000, 64), (100, 20, 64 * 3)):
        stripes.append((rs.New(k, m, rs.WithLeopardGF16(True)), k, m, _leo_codeword(rng, 16, k, m, size)))
    b = rs.Batch(0)
    for use_graph in (True, False):
        b.set_graph(use_graph)
        items = []
        for enc, k, m, full in stripes:
            items.append((enc, rs.OP_ENCODE, [torch.from_numpy(x).to("cuda") for x in full[:k]] +
                          [torch.zeros(len(full[0]), dtype=torch.uint8, device="cuda") for _ in range(m)]))
        enc, k, m, full = stripes[0]
        lost = set(rng.choice(k + m, m, replace=False).tolist())
        rec = [None if i in lost else torch.from_numpy(full[i]).to("cuda") for i in range(k + m)]
        items.append((enc, rs.OP_RECONSTRUCT, rec))
        items.append((enc, rs.OP_VERIFY, [torch.from_numpy(x).to("cuda") for x in full]))
        bad = [torch.from_numpy(x).to("cuda") for x in full]
        bad[k + m - 1][len(full[0]) - 1] ^= 0x10
        items.append((enc, rs.OP_VERIFY, bad))
        st = b.run(items)
        assert st == [0] * len(items), st
        assert b.verify_result(len(items) - 2) and not b.verify_result(len(items) - 1)
        for (enc_, k_, m_, full_), (_, _, sh) in zip(stripes, items):
            for i in range(k_, k_ + m_):
                assert np.array_equal(sh[i].cpu().numpy(), full_[i]), (use_graph, k_, m_, i)
        for i in lost:
            assert np.array_equal(rec[i].cpu().numpy(), full[i]), (use_graph, i)


def test_batched_leopard_unsupported_items_are_reported():
    rng = np.random.default_rng(9)
    # GF(2^16) reconstruct on shards that are not 16-byte aligned.
    k, m, size = 100, 20, 128
    enc = rs.New(k, m, rs.WithLeopardGF16(True))
    full = _leo_codeword(rng, 16, k, m, size)
    base16 = torch.zeros((k + m) * size + 1, dtype=torch.uint8, device="cuda")
    work = [None if i < 3 else base16[1 + i * size: 1 + (i + 1) * size] for i in range(k + m)]
    for i in range(3, k + m):
        work[i].copy_(torch.from_numpy(full[i]))
    # unaligned GF(2^8) encode
    enc8 = rs.New(4, 2, rs.WithLeopardGF(True))
    pool = torch.zeros(6 * 64 + 1, dtype=torch.uint8, device="cuda")
    unaligned = [pool[1 + 64 * i: 1 + 64 * (i + 1)] for i in range(6)]
    short = [torch.zeros(100, dtype=torch.uint8, device="cuda") for _ in range(6)]
    b = rs.Batch(0)
    st = b.run([(enc, rs.OP_RECONSTRUCT, work), (enc8, rs.OP_ENCODE, unaligned), (enc8, rs.OP_ENCODE, short)])
    assert st == [rs.ErrNotSupported.code, rs.ErrNotSupported.code, rs.ErrInvalidShardSize.code]
