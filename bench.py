#!/usr/bin/env python3
"""Benchmark of the plain BSGS sieve on one MI355X.

Workload (R745-shaped, the reference's headline case): base 745, 22 odd k
drawn in [3, 10^7] by seed, n in [180000, 250000] kept when no odd prime
q < 10^5 divides k*745^n-1, and p = every prime in [1e12, 1e12 + K*2^30).
One step = one 2^30-wide p-chunk (about 3.9e7 primes) through the device
sieve (one persistent kernel: set-up and BSGS phases); the prime offsets are resident
in HBM before timing starts (host prime generation is outside the timed
region and reported separately).

value = p-range advanced per second (the reference's "p/sec",
kernel_v0.24.6.cu:1552-1559, :1759), the unit of its published 360e6 p/s
on an A100 (README.md:20); primes/s is reported beside it.

roofline: the BSGS kernel is bound by VALU 32-bit integer multiply-adds
(neither HBM nor MFMA).  achieved = algorithmic v_mad_u64_u32-class
multiplies per launch / mean launch time (HIP events on the sieve's own
stream), peak = the same instruction's rate measured on this device.
cpu_baseline: the CPU oracle (a C++ restatement of the reference
algorithm, OpenMP over primes) on a bounded sample of chunk 0.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "cuda-riesel-sieve_amd"), os.path.join(ROOT, "oracle")]

A100_PUBLISHED = 360e6  # README.md:20, -b 5 -m 8 -s 192 -Q 18, R745.ABCD


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=0,
                    help="baby steps (0 = autotuned, or the cost model with --no-tune)")
    ap.add_argument("--tile", type=int, default=0, help="giant-step tile (0 = autotuned / 2)")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--chunk-log2", type=int, default=30)
    ap.add_argument("--max-batch", type=int, default=1 << 22,
                    help="primes per host-to-device copy (host submits only)")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(p-chunks shard across GPUs with no exchange; see the sharding design)",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np

    import pyrsieve as rs

    if rs.device_count() < 1:
        log("bench.py: no HIP device visible")
        sys.exit(1)
    K, W = args.steps, args.warmup
    chunk = 1 << args.chunk_log2
    p0 = 10**12
    t = time.time()
    abcd = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=args.seed)
    plan = rs.Plan(abcd)
    log(f"job: Q={plan.Q} rows={plan.nrows} n=[{plan.minN},{plan.maxN}] "
        f"({time.time() - t:.1f}s)")
    t = time.time()
    gen = rs.PrimeGen(p0 + K * chunk + 1, threads=min(16, os.cpu_count() or 1))
    bufs, counts = [], []
    for i in range(K):
        lo = p0 + i * chunk
        offs = gen.range(lo, lo + chunk)
        counts.append(len(offs))
        bufs.append(rs.DeviceBuffer(0, offs))
        if i == 0:
            sample_offs = offs
    gen_s = time.time() - t
    nprimes = sum(counts)
    log(f"primes: {nprimes} in [{p0}, {p0 + K * chunk}) generated+uploaded in {gen_s:.1f}s")
    tune = None
    m, tile = args.m, args.tile
    if not args.no_tune and m == 0:
        ts = time.time()
        m, tile, cands = rs.autotune(plan, p0, bufs[0], count=min(counts[0], 1 << 22))
        tune = {"sample_primes": min(counts[0], 1 << 22), "best_m": m, "best_tile": tile,
                "candidates": [[cm, ct, round(ns, 4)] for cm, ct, ns in cands]}
        log(f"autotune: m={m} tile={tile} ({len(cands)} candidates, {time.time() - ts:.1f}s)")
    sv = rs.Sieve(plan, m=m, tile=tile, max_batch=args.max_batch, hit_capacity=1 << 22)
    log(f"sieve: m={sv.m} tile={sv.tile} T={sv.T} slots={sv.slots} lds/wave={sv.lds_bytes}B")
    for w in range(W):
        sv.submit_device(p0 + (w % K) * chunk, bufs[w % K])
        sv.collect()
        log(f"warmup step {w} done")
    sv.timing()  # reset the event accumulators
    t0 = time.perf_counter()
    for i in range(K):
        sv.submit_device(p0 + i * chunk, bufs[i])
    hits = sv.collect()
    t1 = time.perf_counter()
    tm = sv.timing()
    elapsed = t1 - t0
    log(f"timed: {K} steps in {elapsed:.3f}s, {len(hits)} factors, "
        f"kernel {tm['kernel_ms']:.1f} ms")
    prange_rate = K * chunk / elapsed
    primes_rate = nprimes / elapsed

    # algorithmic multiplies per prime in the sieve kernel's BSGS phase: m
    # baby steps and Q*T Y products at 11 32-bit multiplies each (full 64x64
    # Montgomery), A*T giant steps at 5 (32x64 Montgomery), A = rows/2
    # expected Jacobi survivors (the set-up phase is not counted)
    A = plan.nrows / 2.0
    muls_per_prime = 11 * sv.m + 11 * plan.Q * sv.T + 5 * A * sv.T
    launches = K  # one persistent launch per chunk
    per_launch_primes = nprimes / launches
    bsgs_s_per_launch = tm["kernel_ms"] / 1e3 / launches
    achieved = muls_per_prime * per_launch_primes / bsgs_s_per_launch
    peak = rs.dev_mad_rate()
    # streaming copy of a known size: HBM reference for the traffic counters
    copy_bytes = 1 << 30
    hbm = rs.dev_copy_rate(copy_bytes)
    # HBM bytes per prime from the committed counter summary (FETCH_SIZE and
    # WRITE_SIZE scaled by the copy-probe calibration), times primes per launch
    traffic = None
    prof = os.path.join(ROOT, "profiles", "rs_sieve_kernel_r745_pmc.txt")
    if os.path.exists(prof):
        for line in open(prof):
            if line.startswith("traffic_bytes_per_prime:"):
                traffic = float(line.split(":")[1]) * per_launch_primes
    roof = {"bound": "valu", "achieved": achieved / 1e9, "peak": peak / 1e9, "unit": "Gmul/s",
            "frac": achieved / peak, "traffic": traffic,
            "traffic_source": "profiles/rs_sieve_kernel_r745_pmc.txt (m=576, tile=2)",
            "kernel": "rs_sieve_kernel", "muls_per_prime": muls_per_prime,
            "primes_per_launch": per_launch_primes,
            "mean_launch_ms": bsgs_s_per_launch * 1e3,
            "hbm_copy_GBps": hbm / 1e9, "copy_probe_bytes": copy_bytes}

    # end to end on the device: primes generated on the GPU, then sieved
    # (same p-range, no host generation, no PCIe), as an extra figure
    sv.collect()
    sv.timing()
    t0 = time.perf_counter()
    for i in range(K):
        sv.submit_range(p0 + i * chunk, p0 + (i + 1) * chunk)
    hits_e2e = sv.collect()
    e2e_s = time.perf_counter() - t0
    e2e_primes = sv.timing()["primes"]
    if hits_e2e != hits or e2e_primes != nprimes:
        raise RuntimeError("device-generated primes disagree with the host-generated run")
    end_to_end = {"value": K * chunk / e2e_s, "unit": "p/sec", "primes_per_sec": e2e_primes / e2e_s,
                  "what": "GPU prime generation + sieve of the same p-range, timed together"}
    log(f"end to end (GPU primes): {K * chunk / e2e_s:.3e} p/s")

    cpu = None
    if not args.no_cpu:
        import oracle

        rk, rr, _ = plan.rows()
        job = oracle.Job(plan.base, plan.Q, sv.m, plan.minN, plan.maxN, rk, rr)
        n = 2000
        while True:
            ts = time.perf_counter()
            oracle.bsgs_many(job, p0, sample_offs[:n], threads=args.cpu_threads)
            dt = time.perf_counter() - ts
            if dt > args.cpu_seconds / 4 or n >= len(sample_offs):
                break
            n = min(len(sample_offs), int(n * max(2.0, args.cpu_seconds / 4 / max(dt, 1e-3))))
        span = int(sample_offs[n - 1]) + 1
        cpu = {"value": span / dt, "unit": "p/sec", "cores": args.cpu_threads, "kind": "port",
               "primes_per_sec": n / dt,
               "sample": f"first {n} primes of [1e12, 1e12+2^30) (p-range {span}), "
                         f"same job and m, {dt:.2f}s"}
        log(f"cpu: {n} primes in {dt:.2f}s on {args.cpu_threads} threads")

    out = {
        "metric": "primes/sec sieved (22 k, n 180k-250k) plain vs batched-serve; p99 job latency",
        "value": prange_rate,
        "unit": "p/sec (p-range advanced per second, the reference's unit)",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": elapsed / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": prange_rate / A100_PUBLISHED,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "R745-shaped plain BSGS: base 745, 22 odd k in [3,1e7] "
                               f"(seed {args.seed}), n in [180000,250000] presieved by odd "
                               f"primes < 1e5, p in [1e12, 1e12+{K}*2^{args.chunk_log2})",
                   "Q": plan.Q, "rows": plan.nrows, "m": sv.m, "T": sv.T, "tile": sv.tile,
                   "primes": nprimes, "chunk": chunk, "global_batch": counts[0],
                   "parallelism": "single GPU, wavefront per prime"},
        "primes_per_sec": primes_rate,
        "factors": len(hits),
        "kernel_ms": tm["kernel_ms"],
        "end_to_end_gpu_primes": end_to_end,
        "host_primegen_s": gen_s,
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotune": tune,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
