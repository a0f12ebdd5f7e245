"""GPU tests of batched Verify items and of the request loop (continuous
batching under an HBM budget): every request's result in a served mix must
equal the same request run alone through the plain per-stripe method, byte
for byte, with the same status and Verify outcome."""
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


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def _codeword(o, k, m, size, rng):
    data = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)]
    ref = data + [np.zeros(size, np.uint8) for _ in range(m)]
    assert o.encode(ref) == 0
    return ref


def test_batch_verify_items_match_plain_verify():
    rng = np.random.default_rng(11)
    cases = []
    for k, m, size in [(10, 4, 65536), (6, 3, 1000), (17, 3, 4097), (4, 2, 31), (5, 0, 512)]:
        enc = rs.New(k, m)
        o = rsref.Classic(k, m)
        for corrupt in (False, True):
            if corrupt and m == 0:
                continue
            ref = _codeword(o, k, m, size, rng)
            shards = [_dev(r) for r in ref]
            if corrupt:
                i = int(rng.integers(0, k + m))
                shards[i][int(rng.integers(0, size))] ^= 0x5A
            cases.append((enc, shards, not corrupt))
    b = rs.Batch(0)
    items = [(enc, rs.OP_VERIFY, shards) for enc, shards, _ in cases]
    for use_graph in (True, False):
        b.set_graph(use_graph)
        assert b.run(items) == [0] * len(items)
        for i, (enc, shards, want) in enumerate(cases):
            assert b.verify_result(i) == want, (i, use_graph)
            assert enc.Verify(shards) == want
    # a verify item beside encode items, and an invalid verify item
    enc = rs.New(10, 4)
    ref = _codeword(rsref.Classic(10, 4), 10, 4, 4096, rng)
    good = [_dev(r) for r in ref]
    fresh = [_dev(r) for r in ref[:10]] + [torch.zeros(4096, dtype=torch.uint8, device="cuda") for _ in range(4)]
    short = [_dev(r) for r in ref]
    short[2] = short[2][:100]
    st = b.run([(enc, rs.OP_ENCODE, fresh), (enc, rs.OP_VERIFY, good), (enc, rs.OP_VERIFY, short)])
    assert st == [0, 0, rs.ErrShardSize.code]
    assert b.verify_result(1)
    for i in range(10, 14):
        assert np.array_equal(fresh[i].cpu().numpy(), ref[i])


def test_batch_rejects_host_and_missing_buffers():
    enc = rs.New(4, 2)
    host = [torch.zeros(256, dtype=torch.uint8) for _ in range(6)]
    mixed = [torch.zeros(256, dtype=torch.uint8, device="cuda") for _ in range(5)] + [host[0]]
    b = rs.Batch(0)
    st = b.run([(enc, rs.OP_ENCODE, host), (enc, rs.OP_ENCODE, mixed)])
    assert st == [rs.ErrMixedMemory.code, rs.ErrMixedMemory.code]


# --- served mix ----------------------------------------------------------

CLASSIC = [(10, 4), (6, 3), (16, 4), (4, 2), (17, 3), (20, 12)]
SIZES = [1000, 4096, 65536, 1 << 18, 1 << 20]


def _make_requests(rng, n, leopard_every=0, leo_shape=(8, 192, 56, 65536)):
    """Returns [(encoder, op, shards_solo, shards_mix, note)] with identical
    independent inputs for the solo and the mixed run."""
    codecs = {s: rs.New(*s) for s in CLASSIC}
    oracles = {s: rsref.Classic(*s) for s in CLASSIC}
    lbits, lk, lm, lsize = leo_shape
    lopt = rs.WithLeopardGF16(True) if lbits == 16 else rs.WithLeopardGF(True)
    leo = rs.New(lk, lm, lopt) if leopard_every else None
    leo_o = rsref.Leopard(lk, lm, bits=lbits) if leopard_every else None
    out = []
    for n_ in range(n):
        if leopard_every and n_ % leopard_every == leopard_every - 1:
            k, m, size, enc, o = lk, lm, lsize, leo, leo_o
        else:
            k, m = CLASSIC[int(rng.integers(0, len(CLASSIC)))]
            size = int(rng.choice(SIZES))
            enc, o = codecs[(k, m)], oracles[(k, m)]
        ref = _codeword(o, k, m, size, rng)
        r = rng.random()
        if r < 0.6:
            op = rs.OP_ENCODE
            host = ref[:k] + [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(m)]  # stale parity
            lost = set()
        elif r < 0.7:
            op = rs.OP_VERIFY
            host = [x.copy() for x in ref]
            if rng.random() < 0.5:
                host[int(rng.integers(0, k + m))][int(rng.integers(0, size))] ^= 1
            lost = set()
        else:
            op = rs.OP_RECONSTRUCT if r < 0.85 else rs.OP_RECONSTRUCT_DATA
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            host = ref
        solo = [None if i in lost else _dev(host[i]) for i in range(k + m)]
        mix = [None if i in lost else _dev(host[i]) for i in range(k + m)]
        out.append((enc, op, solo, mix, (k, m, size, ref, lost)))
    return out


def _run_solo(enc, op, shards):
    if op == rs.OP_ENCODE:
        enc.Encode(shards)
        return None
    if op == rs.OP_VERIFY:
        return enc.Verify(shards)
    if op == rs.OP_RECONSTRUCT:
        enc.Reconstruct(shards)
    else:
        enc.ReconstructData(shards)
    return None


def _check_same(req, vok, solo_ok):
    enc, op, solo, mix, (k, m, size, ref, lost) = req
    if op == rs.OP_VERIFY:
        assert vok == solo_ok, (k, m, size)
        return
    for i in range(k + m):
        a, b = solo[i], mix[i]
        assert (a is None) == (b is None), (op, k, m, i)
        if a is not None:
            assert torch.equal(a, b), (op, k, m, size, i)
    # and both hold the codeword
    want = range(k + m) if op != rs.OP_RECONSTRUCT_DATA else range(k)
    for i in want:
        assert np.array_equal(mix[i].cpu().numpy(), ref[i]), (op, k, m, size, i)


# GF(2^8) (192, 56) 64 KiB as in the served stream; GF(2^16) (5000, 100)
# reconstructs n = 8192 work rows (the plan's three-pass launches).
@pytest.mark.parametrize("leo_shape", [(8, 192, 56, 65536), (16, 5000, 100, 128)])
def test_served_mix_equals_solo_runs(leo_shape):
    rng = np.random.default_rng(5150)
    reqs = _make_requests(rng, 96, leopard_every=12, leo_shape=leo_shape)
    solo_ok = [_run_solo(enc, op, solo) for enc, op, solo, _, _ in reqs]
    torch.cuda.synchronize()
    srv = rs.Server(0, hbm_budget=256 << 20)
    gaps = rng.exponential(20_000, len(reqs))  # mean 20 us between arrivals
    arrivals = np.cumsum(gaps).astype(np.uint64)
    status, vok, lat, span = srv.run_trace([(e, op, mix) for e, op, _, mix, _ in reqs], arrivals)
    assert status == [0] * len(reqs)
    assert span > 0 and all(x > 0 for x in lat)
    for req, v, s in zip(reqs, vok, solo_ok):
        _check_same(req, v, s)
    batches, batched, max_inflight = srv.stats()
    assert batches >= 1 and batched <= len(reqs)
    srv.close()


def test_server_submit_wait_budget_and_errors():
    rng = np.random.default_rng(99)
    reqs = _make_requests(rng, 40)
    solo_ok = [_run_solo(enc, op, solo) for enc, op, solo, _, _ in reqs]
    biggest = max(r[4][2] * (r[4][0] + r[4][1]) for r in reqs)
    budget = biggest  # a request larger than the budget is rejected at submit
    srv = rs.Server(0, hbm_budget=budget, max_batch=8)
    tickets = [srv.submit(e, op, mix) for e, op, _, mix, _ in reqs]
    enc = rs.New(5, 2)
    bad = [torch.zeros(256, dtype=torch.uint8, device="cuda") for _ in range(7)]
    bad[3] = bad[3][:100]
    t_bad = srv.submit(enc, rs.OP_ENCODE, bad)
    for req, t, s in zip(reqs, tickets, solo_ok):
        st, v, lat = srv.wait(t)
        assert st == 0
        _check_same(req, v, s)
    assert srv.wait(t_bad)[0] == rs.ErrShardSize.code
    srv.drain()
    _, _, max_inflight = srv.stats()
    assert 0 < max_inflight <= budget
    with pytest.raises(rs.RSError):
        srv.submit(enc, rs.OP_ENCODE, bad[:6])
    srv.close()


def _encode_request(rng, k, m, size):
    """An Encode request with stale parity: (encoder, shards, codeword)."""
    ref = _codeword(rsref.Classic(k, m), k, m, size, rng)
    host = ref[:k] + [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(m)]
    return rs.New(k, m), [_dev(h) for h in host], ref, host


def test_server_cancel_pause_and_admission_rejection():
    rng = np.random.default_rng(2323)
    srv = rs.Server(0, hbm_budget=8 << 20, max_batch=8)
    srv.pause()
    reqs = [_encode_request(rng, 10, 4, 65536) for _ in range(6)]
    tickets = [srv.submit(enc, rs.OP_ENCODE, sh) for enc, sh, _, _ in reqs]
    assert srv.cancel(tickets[1]) and srv.cancel(tickets[4])
    assert not srv.cancel(tickets[4])  # already complete (cancelled), not yet reported
    per = 14 * 65536
    assert srv.admission() == (0, 2, 4 * per)
    # a request larger than the whole budget is rejected at submit
    big_enc, big, _, _ = _encode_request(rng, 10, 4, 1 << 20)
    with pytest.raises(rs.ErrOverBudget):
        srv.submit(big_enc, rs.OP_ENCODE, big)
    # the queue limit rejects what would take the queued bytes past it
    srv.set_queue_limit(5 * per)
    extra = _encode_request(rng, 10, 4, 65536)
    t_extra = srv.submit(extra[0], rs.OP_ENCODE, extra[1])
    with pytest.raises(rs.ErrOverBudget):
        srv.submit(extra[0], rs.OP_ENCODE, _encode_request(rng, 10, 4, 65536)[1])
    assert srv.admission() == (2, 2, 5 * per)
    with pytest.raises(rs.ErrInvalidInput):
        srv.drain()  # paused with requests queued: would never finish
    srv.set_queue_limit(0)
    srv.pause(False)
    for i, (t, (enc, sh, ref, host)) in enumerate(zip(tickets, reqs)):
        st, _, _ = srv.wait(t)
        if i in (1, 4):
            assert st == rs.ErrCancelled.code
            for j in range(14):  # a cancelled request never ran: stale parity stays
                assert np.array_equal(sh[j].cpu().numpy(), host[j])
        else:
            assert st == 0
            for j in range(14):
                assert np.array_equal(sh[j].cpu().numpy(), ref[j]), (i, j)
    assert srv.wait(t_extra)[0] == 0
    # a finished request cannot be cancelled; a reported ticket is unknown
    enc, sh, ref, _ = _encode_request(rng, 6, 3, 4096)
    t = srv.submit(enc, rs.OP_ENCODE, sh)
    srv.drain()
    assert not srv.cancel(t)
    assert srv.wait(t)[0] == 0
    with pytest.raises(rs.ErrInvalidInput):
        srv.cancel(t)
    # a trace with an oversized request: its status is the rejection
    trace = [_encode_request(rng, 4, 2, 4096) for _ in range(3)]
    trace.insert(1, _encode_request(rng, 10, 4, 1 << 20))
    st, _, lat, _ = srv.run_trace([(e, rs.OP_ENCODE, sh) for e, sh, _, _ in trace], [0, 1000, 2000, 3000])
    assert st == [0, rs.ErrOverBudget.code, 0, 0] and lat[1] == 0
    for i in (0, 2, 3):
        e, sh, ref, _ = trace[i]
        for j in range(6):
            assert np.array_equal(sh[j].cpu().numpy(), ref[j])
    srv.close()
