"""Throughput of the MI355X erasure-coding hot path on one GPU.

Workload (the plain-GPU RS(16,4) case): 256 stripes of 16 data shards x
4 MiB of seeded uniform random bytes plus 4 parity shards, resident in HBM.
One step = Encode every stripe, then Reconstruct every stripe after erasing
a seeded random set of 0-4 of its 20 shards (data or parity).  The erased
shards are rebuilt in place, so every step sees the same state.

Bytes per operation follow the reference's benchmarks
(b.SetBytes(shardSize*(dataShards+parityShards))): (k+m)*shard for each
Encode and for each Reconstruct that has work to do (a stripe with no
erasure returns immediately and is not counted).

    python bench.py [--steps K] [--warmup W] [--stripes N] [--shard-mib S]

Prints one JSON line.  `roofline` is for the dominant kernel (the encode
pass, 16 inputs -> 4 outputs), timed per launch with HIP events on the
codec's stream; `cpu_baseline` times the CPU oracle on one stripe.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E, 8 TB/s
METRIC = "encode+reconstruct GB/s and % HBM peak, batched vs plain GPU; p50/p99 latency"


def _parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--stripes", type=int, default=256)
    ap.add_argument("--shard-mib", type=float, default=4.0)
    ap.add_argument("--k", type=int, default=16)
    ap.add_argument("--m", type=int, default=4)
    ap.add_argument("--seed", type=int, default=0xABADC0CAC01A)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    return ap.parse_args()


def _log(msg):
    print(msg, file=sys.stderr, flush=True)


def main() -> int:
    args = _parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is described in DESIGN.md, not launched)"}))
        return 2
    import torch
    import reedsolomon_amd as rs

    k, m = args.k, args.m
    n = k + m
    S = int(args.shard_mib * (1 << 20))
    S -= S % 64
    nst = args.stripes
    _log(f"bench: RS({k},{m}) stripes={nst} shard={S} B, allocating {nst * n * S / 2**30:.1f} GiB")

    torch.cuda.set_device(0)
    enc = rs.New(k, m, rs.WithMI355X(0))
    h = enc.handle
    lib = rs.lib
    stream_ptr = lib.rsx_get_stream(h)

    gen = torch.Generator(device="cuda")
    gen.manual_seed(args.seed & 0x7FFFFFFFFFFFFFFF)
    buf = torch.empty((nst, n, S), dtype=torch.uint8, device="cuda")
    buf[:, :k].random_(0, 256, generator=gen)
    buf[:, k:].zero_()
    torch.cuda.synchronize()
    base = buf.data_ptr()

    # Per-stripe pointer/length arrays for the C ABI.
    PtrArr = ctypes.c_void_p * n
    LenArr = ctypes.c_size_t * n
    ptrs = []
    full_lens = LenArr(*([S] * n))
    for s in range(nst):
        ptrs.append(PtrArr(*[base + (s * n + i) * S for i in range(n)]))
    rng = np.random.default_rng(args.seed)
    erase_lens = []
    rec_bytes = 0
    for s in range(nst):
        e = int(rng.integers(0, m + 1))
        lost = rng.choice(n, e, replace=False) if e else []
        lens = LenArr(*[0 if i in set(np.asarray(lost).tolist()) else S for i in range(n)])
        erase_lens.append((lens, e))
        if e:
            rec_bytes += n * S
    filled = (ctypes.c_uint8 * n)()

    def check(rc):
        if rc != 0:
            raise RuntimeError(f"rsx error {rc}: {lib.rsx_last_error().decode()}")

    def step():
        for s in range(nst):
            check(lib.rsx_encode(h, ptrs[s], full_lens, n))
        for s in range(nst):
            lens, e = erase_lens[s]
            if e:
                check(lib.rsx_reconstruct(h, ptrs[s], lens, n, filled))

    # Parity check of one column slice of stripe 0 against the CPU oracle
    # (encode is column-separable, so a slice is a complete sub-problem).
    for _ in range(max(args.warmup, 1)):
        step()
    check(lib.rsx_synchronize(h))
    parity_ok = None
    try:
        from oracle import rsref
        sl = slice(0, 1 << 16)
        stripe0 = buf[0, :, sl].cpu().numpy()
        o = rsref.Classic(k, m)
        ref = [stripe0[i].copy() for i in range(k)] + [np.zeros(sl.stop, np.uint8) for _ in range(m)]
        o.encode(ref)
        parity_ok = all(np.array_equal(ref[k + i], stripe0[k + i]) for i in range(m))
    except Exception as ex:  # the oracle is a checker only
        _log(f"bench: oracle check skipped: {ex}")

    # Timed region: K steps, bracketed by device synchronisation.
    torch.cuda.synchronize()
    check(lib.rsx_synchronize(h))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    check(lib.rsx_synchronize(h))
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    bytes_per_step = nst * n * S + rec_bytes
    gbps = bytes_per_step * args.steps / elapsed / 1e9
    _log(f"bench: {args.steps} steps in {elapsed:.3f}s -> {gbps:.1f} GB/s")

    # Dominant kernel: one encode launch per call (16 in -> 4 out).  Time each
    # launch with HIP events on the codec's stream.
    stream = torch.cuda.ExternalStream(stream_ptr)
    nlaunch = min(nst, 128)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(nlaunch)]
    for s in range(nlaunch):
        evs[s][0].record(stream)
        check(lib.rsx_encode(h, ptrs[s], full_lens, n))
        evs[s][1].record(stream)
    check(lib.rsx_synchronize(h))
    durs = [a.elapsed_time(b) * 1e-3 for a, b in evs]
    avg_launch = float(np.mean(durs))
    enc_bytes = n * S
    achieved = enc_bytes / avg_launch / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": f"gf8_pass_vec16<{m}> (RS({k},{m}) encode, {enc_bytes} B per launch)",
                "avg_launch_us": round(avg_launch * 1e6, 2)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(k, m, S, args.cpu_seconds)

    out = {
        "metric": METRIC,
        "value": round(gbps, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {
            "workload": f"plain GPU RS({k},{m}) GF(2^8): {nst} stripes x {k}x{S // (1 << 20)} MiB "
                        f"data shards; encode all, then reconstruct each after 0-{m} seeded erasures",
            "k": k, "m": m, "shard_bytes": S, "stripes": nst,
            "path": "plain (one C-ABI call per stripe, device-resident shards)",
            "bytes_per_step": bytes_per_step,
        },
        "pct_hbm_peak": round(100.0 * gbps / HBM_PEAK_GBPS, 2),
        "parity_check_vs_oracle": parity_ok,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


def cpu_baseline(k: int, m: int, S: int, seconds: float):
    """The CPU oracle (scalar restatement, one thread) on a bounded sample:
    repeated RS(k,m) encodes of one stripe of 1 MiB shards."""
    try:
        from oracle import rsref
    except Exception as ex:
        _log(f"bench: cpu baseline skipped: {ex}")
        return None
    size = min(S, 1 << 20)
    rng = np.random.default_rng(1)
    shards = [rng.integers(0, 256, size, dtype=np.uint8) for _ in range(k)] + \
             [np.zeros(size, np.uint8) for _ in range(m)]
    o = rsref.Classic(k, m)
    t0 = time.perf_counter()
    reps = 0
    while True:
        o.encode(shards)
        reps += 1
        el = time.perf_counter() - t0
        if el >= seconds:
            break
    gbps = reps * (k + m) * size / el / 1e9
    return {"value": round(gbps, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{reps} RS({k},{m}) encodes of one stripe of {size} B shards "
                      f"({el:.1f} s, scalar oracle, 1 thread)"}


if __name__ == "__main__":
    sys.exit(main())
