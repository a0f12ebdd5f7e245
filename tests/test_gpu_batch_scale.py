"""The batched path at the configured workload scales, against the oracle:
RS(16,4) over 4 MiB shards (encode plan, then a reconstruct plan after 0-4
seeded erasures per stripe), and a mixed plan of 1024 stripes of the five
mixed-workload shapes with shard sizes up to 1 MiB (30% reconstructs), with
the tile -> task table and with the binary search that replaces it on very
large plans.  Reconstructed shards must equal the original bytes; parity is
compared with the oracle on column slices (every byte column is computed
independently, so a slice is a complete check of those columns)."""
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


def _oracle_parity(k, m, data_cols):
    ref = [np.ascontiguousarray(d) for d in data_cols] + [np.zeros(len(data_cols[0]), np.uint8) for _ in range(m)]
    assert rsref.Classic(k, m).encode(ref) == 0
    return ref[k:]


def test_rs16_4_4mib_batched_vs_oracle():
    k, m, size, nst = 16, 4, 4 << 20, 8
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1604)
    buf = torch.zeros((nst, k + m, size), dtype=torch.uint8, device="cuda")
    buf[:, :k].random_(0, 256, generator=gen)
    enc = rs.New(k, m)
    b = rs.Batch(0)
    st = b.run([(enc, rs.OP_ENCODE, [buf[s, i] for i in range(k + m)]) for s in range(nst)])
    assert st == [0] * nst
    # parity against the oracle on three 64 KiB column slices of every stripe
    for s in range(nst):
        for c0 in (0, size // 2 - 12345, size - (64 << 10)):
            cols = buf[s, :, c0:c0 + (64 << 10)].cpu().numpy()
            want = _oracle_parity(k, m, list(cols[:k]))
            for i in range(m):
                assert np.array_equal(cols[k + i], want[i]), (s, c0, i)
    orig = buf.clone()
    rng = np.random.default_rng(1604)
    items, lost_sets = [], []
    for s in range(nst):
        e = int(rng.integers(0, m + 1)) if s else m  # stripe 0: the most erasures
        lost = set(rng.choice(k + m, e, replace=False).tolist())
        for i in lost:
            buf[s, i].zero_()
        lost_sets.append(lost)
        items.append((enc, rs.OP_RECONSTRUCT, [None if i in lost else buf[s, i] for i in range(k + m)]))
    st = rs.Batch(0).run(items)
    assert st == [0] * nst
    for s, (_, _, shards) in enumerate(items):
        for i in lost_sets[s]:
            assert torch.equal(shards[i], orig[s, i]), (s, i)


def _mixed_plan(env_max, monkeypatch):
    if env_max is None:
        monkeypatch.delenv("RSX_BATCH_TILE_TABLE_MAX", raising=False)
    else:
        monkeypatch.setenv("RSX_BATCH_TILE_TABLE_MAX", str(env_max))
    shapes_all = [(4, 2), (6, 3), (10, 4), (17, 3), (32, 8)]
    rng = np.random.default_rng(3003)
    nst = 1024
    shapes = [shapes_all[int(i)] for i in rng.integers(0, 5, nst)]
    sizes = [max(64, int(round(np.exp(rng.uniform(np.log(4096), np.log(1 << 20))) / 64)) * 64) for _ in range(nst)]
    offs, total = [], 0
    for (k, m), s in zip(shapes, sizes):
        offs.append(total)
        total += (k + m) * s
    buf = torch.zeros(total, dtype=torch.uint8, device="cuda")
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3003)
    for (k, m), s, o in zip(shapes, sizes, offs):
        buf[o:o + k * s].random_(0, 256, generator=gen)
    codecs = {sh: rs.New(*sh) for sh in shapes_all}

    def shard(i, j):
        s, o = sizes[i], offs[i]
        return buf[o + j * s:o + (j + 1) * s]

    def stripe(i):
        k, m = shapes[i]
        return [shard(i, j) for j in range(k + m)]

    st = rs.Batch(0).run([(codecs[shapes[i]], rs.OP_ENCODE, stripe(i)) for i in range(nst)])
    assert st == [0] * nst
    orig = buf.clone()
    items, lost_sets = [], []
    for i in range(nst):
        k, m = shapes[i]
        if rng.random() < 0.3:
            lost = set(rng.choice(k + m, int(rng.integers(1, m + 1)), replace=False).tolist())
            for j in lost:
                shard(i, j).zero_()
            items.append((codecs[shapes[i]], rs.OP_RECONSTRUCT, [None if j in lost else shard(i, j) for j in range(k + m)]))
        else:
            lost = set()
            for j in range(k, k + m):
                shard(i, j).zero_()  # the encode rewrites it
            items.append((codecs[shapes[i]], rs.OP_ENCODE, stripe(i)))
        lost_sets.append(lost)
    b = rs.Batch(0)
    st = b.plan(items)
    assert st == [0] * nst
    b.launch()
    b.synchronize()
    # every reconstruct rebuilt its lost shards (into output tensors the plan
    # gave them) and every encode rewrote its parity
    for i, (_, op, shards) in enumerate(items):
        k, m = shapes[i]
        for j in lost_sets[i]:
            want = orig[offs[i] + j * sizes[i]:offs[i] + (j + 1) * sizes[i]]
            assert torch.equal(shards[j], want), (i, j)
            shard(i, j).copy_(shards[j])
    assert torch.equal(buf, orig)
    # the first encode's parity (hence all of it) against the oracle on a sample
    for i in rng.choice(nst, 48, replace=False):
        k, m = shapes[i]
        c = min(sizes[i], 8192)
        cols = [shard(i, j)[:c].cpu().numpy() for j in range(k + m)]
        want = _oracle_parity(k, m, cols[:k])
        for j in range(m):
            assert np.array_equal(cols[k + j], want[j]), (i, j)
    return b


def test_mixed_1024_stripes_batched_vs_oracle(monkeypatch):
    _mixed_plan(None, monkeypatch)


def test_mixed_plan_binary_search_tile_lookup(monkeypatch):
    # RSX_BATCH_TILE_TABLE_MAX=0: blocks find their task by binary search over
    # the tasks' first tiles (the path of plans above 2^24 tiles).
    _mixed_plan(0, monkeypatch)
