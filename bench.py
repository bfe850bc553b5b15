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

batched: 64 heterogeneous jobs (workloads.job_specs: base U[3,1030], k-count
log-U[1,200], n-width log-U[1e4,5e5], presieve < 1e5) over the p-chunk
[1e12, 1e12+2^30) with per-job autotuned (m, tile): the jobs run plain
back-to-back and then as one graph-captured batch; per-job hit sets must be
equal.
serving: jobs of the same distribution each consume consecutive 2^30-wide
p-chunks from 1e12 (the per-job bound is 1e12 + C*2^30, C = --serve-chunks,
scaled down from 1e13 so the run takes seconds); the continuous-batching
server runs them saturated (all arrive at once) and under Poisson arrivals,
reporting p50/p99 job latency and aggregate primes/s against the same jobs
plain back-to-back; per-job hit sets must equal the plain runs.
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
    ap.add_argument("--batch-jobs", type=int, default=64)
    ap.add_argument("--serve-jobs", type=int, default=24)
    ap.add_argument("--serve-chunks", type=int, default=4)
    ap.add_argument("--serve-load", type=float, default=0.7,
                    help="Poisson arrival rate as a fraction of the plain job rate")
    ap.add_argument("--serve-active", type=int, default=16)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-sweep", type=str, default="0.4,0.7,1.0,1.3",
                    help="comma-separated loads (fractions of the plain job rate) of the "
                         "Poisson rate sweep; empty: none")
    ap.add_argument("--serve-quantum-ms", type=float, default=40.0,
                    help="device time per job per server step (0: whole 2^30 windows)")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-real", action="store_true")
    ap.add_argument("--no-edge", action="store_true")
    ap.add_argument("--real-chunks", type=int, default=8)
    ap.add_argument("--edge-chunks", type=int, default=2)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-cli", action="store_true")
    ap.add_argument("--no-msweep", action="store_true")
    ap.add_argument("--edge-m", type=int, default=0, help="edge table size (0 = autotuned)")
    ap.add_argument("--edge-tile", type=int, default=0)
    ap.add_argument("--cli-span-log2", type=int, default=35,
                    help="p-range of the drop-in command-line runs")
    ap.add_argument("--only", choices=["plain", "real", "edge", "batched", "serving", "cli",
                                       "msweep", "cpu"],
                    help="run one leg alone and print its JSON (profiling runs)")
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
    if args.only and args.only not in ("plain", "cpu"):
        leg = {"real": lambda: bench_real_r745(args, rs, chunk),
               "edge": lambda: bench_edge(args, rs, chunk),
               "batched": lambda: bench_batched(args, rs, p0, chunk),
               "serving": lambda: bench_serving(args, rs, p0, chunk),
               "cli": lambda: bench_cli(args),
               "msweep": lambda: bench_msweep(args, rs, p0)}[args.only]
        print(json.dumps({args.only: leg()}), flush=True)
        return
    t = time.time()
    abcd = rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=args.seed)
    plan = rs.Plan(abcd)
    log(f"job: Q={plan.Q} rows={plan.nrows} n=[{plan.minN},{plan.maxN}] "
        f"({time.time() - t:.1f}s)")
    if args.only == "cpu":
        print(json.dumps({"cpu": bench_cpu(args, rs, plan, args.m or 512)}), flush=True)
        return
    t = time.time()
    gen = rs.PrimeGen(p0 + K * chunk + 1, threads=min(16, os.cpu_count() or 1))
    bufs, counts = [], []
    host_offs = []
    for i in range(K):
        lo = p0 + i * chunk
        offs = gen.range(lo, lo + chunk)
        counts.append(len(offs))
        bufs.append(rs.DeviceBuffer(0, offs))
        host_offs.append(offs)
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
    log(f"sieve: m={sv.m} tile={sv.tile} T={sv.T} slots={sv.slots} lds/wave={sv.lds_bytes}B "
        f"kernel={sv.kernel}")
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
    roof = roofline(rs, plan, sv, tm, K, nprimes)

    pcie = end_to_end = None
    if args.only != "plain":
        # the same chunks handed over in host memory (the reference's flow:
        # pinned double-buffered copies overlapped with the sieve)
        sv.collect()
        sv.timing()
        t0 = time.perf_counter()
        for i in range(K):
            sv.submit(p0 + i * chunk, host_offs[i])
        hits_host = sv.collect()
        host_s = time.perf_counter() - t0
        if hits_host != hits:
            raise RuntimeError("host-submitted chunks disagree with the device-resident run")
        pcie = {"value": K * chunk / host_s, "unit": "p/sec", "primes_per_sec": nprimes / host_s,
                "what": "same chunks from host memory through pinned double buffers "
                        f"({args.max_batch} primes per copy), PCIe copies included"}
        log(f"host submits (PCIe included): {K * chunk / host_s:.3e} p/s")
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
        end_to_end = {"value": K * chunk / e2e_s, "unit": "p/sec",
                      "primes_per_sec": e2e_primes / e2e_s,
                      "what": "GPU prime generation + sieve of the same p-range, timed together"}
        log(f"end to end (GPU primes): {K * chunk / e2e_s:.3e} p/s")
    del host_offs
    run_m = sv.m
    sv.close()
    for b in bufs:
        b.free()

    legs = not args.only
    cpu = None if args.no_cpu or not legs else bench_cpu(args, rs, plan, run_m)
    if cpu:
        # the survey's count of Montgomery products per prime (the oracle's
        # restatement of the reference kernel at this m, 11 multiplies each)
        mm = cpu.pop("_montmuls_per_prime")
        roof["survey_montmuls_per_prime"] = mm
        roof["survey_frac_of_published_valu_peak"] = (
            11 * mm * primes_rate / roof["published_valu_peak"] / 1e9)
    real = None if args.no_real or not legs else bench_real_r745(args, rs, chunk)
    cli = None if args.no_cli or not legs else bench_cli(args)
    edge = None if args.no_edge or not legs else bench_edge(args, rs, chunk)
    msweep = None if args.no_msweep or not legs else bench_msweep(args, rs, p0)
    batched = None if args.no_batch or not legs else bench_batched(args, rs, p0, chunk)
    serving = None if args.no_serve or not legs else bench_serving(args, rs, p0, chunk)

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
                   "Q": plan.Q, "rows": plan.nrows, "m": run_m, "T": roof["T"],
                   "tile": roof["tile"], "kernel": roof["kernel"],
                   "primes": nprimes, "chunk": chunk, "global_batch": counts[0],
                   "parallelism": "single GPU, wavefront per prime"},
        "primes_per_sec": primes_rate,
        "factors": len(hits),
        "kernel_ms": tm["kernel_ms"],
        "end_to_end_gpu_primes": end_to_end,
        "host_submit_pcie": pcie,
        "host_primegen_s": gen_s,
        "roofline": roof,
        "cpu_baseline": cpu,
        "autotune": tune,
        "r745_abcd": real,
        "drop_in_cli": cli,
        "edge_base2": edge,
        "m_sweep": msweep,
        "batched": batched,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


def profile_summary(kernel, tile, workload):
    """Counter figures of the committed rocprofv3 summary for this kernel
    build, tile and workload (profiles/<kernel>_t<tile>_<workload>_pmc.txt),
    or {"source": None} when none was made for that build."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_t{tile}_{workload}_pmc.txt")
    out = {"source": None}
    if not os.path.exists(path):
        return out
    out["source"] = os.path.relpath(path, ROOT)
    for line in open(path):
        key, _, val = line.partition(":")
        key = key.strip()
        if key in ("traffic_bytes_per_prime", "valu_insts_per_prime", "valu_issue_busy_frac",
                   "lds_bank_conflict_cycles_per_prime", "salu_insts_per_prime",
                   "lds_active_cycles_per_prime", "lds_busy_frac", "mean_launch_ms"):
            out[key] = float(val.split()[0])
    return out


def roofline(rs, plan, sv, tm, launches, nprimes, workload="r745", probes=True):
    """The dominant kernel's roofline line.  The kernel is bound by VALU
    instruction issue (exact 32-bit integer and FP64 arithmetic) together
    with the LDS (the baby table's random bucket reads), neither HBM (4
    bytes in per prime) nor MFMA.  achieved = algorithmic 32-bit
    multiplies per prime x primes per launch / mean launch time (HIP events
    on the sieve's own stream); peak = the published VALU lane-operation
    rate of MI355X (78.6 TFLOP/s vector FP64 = 39.3e12 full-rate lane
    operations/s: 1024 SIMDs x 16 lanes x 2.4 GHz).  The committed counter
    summary of the same build and tile adds how busy the VALU issue and the
    LDS are, as shares of the CU cycles per prime."""
    # multiplies per prime of the BSGS phase: m baby steps and Q*T Y-table
    # products at 11 (full 64x64 Montgomery), A*T giant steps at 5, A =
    # rows/2 expected Jacobi survivors (the set-up phase is not counted)
    A = plan.nrows / 2.0
    muls = 11 * sv.m + 11 * plan.Q * sv.T + 5 * A * sv.T
    per_launch = nprimes / launches
    launch_s = tm["kernel_ms"] / 1e3 / launches
    achieved = muls * per_launch / launch_s
    published = 78.6e12 / 2
    prof = profile_summary(sv.kernel, sv.tile, workload)
    traffic = prof.get("traffic_bytes_per_prime")
    out = {"bound": "valu", "achieved": achieved / 1e9, "peak": published / 1e9,
           "unit": "Gmul/s", "frac": achieved / published,
           "traffic": traffic * per_launch if traffic else None,
           "traffic_bytes_per_prime": traffic, "profile": prof,
           "kernel": sv.kernel, "m": sv.m, "T": sv.T, "tile": sv.tile,
           "muls_per_prime": muls, "primes_per_launch": per_launch,
           "mean_launch_ms": launch_s * 1e3,
           "published_valu_peak": published / 1e9,
           "hbm_frac": (traffic * per_launch / launch_s / 8e12) if traffic else None,
           "valu_issue_busy_frac": prof.get("valu_issue_busy_frac"),
           "lds_busy_frac": prof.get("lds_busy_frac")}
    if probes:
        mad = rs.dev_mad_rate()
        copy_bytes = 1 << 30
        out.update({"measured_mad_rate": mad / 1e9, "frac_of_measured_mad_rate": achieved / mad,
                    "hbm_copy_GBps": rs.dev_copy_rate(copy_bytes) / 1e9,
                    "copy_probe_bytes": copy_bytes})
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def bench_cpu(args, rs, plan, m):
    """The CPU baseline: the oracle (the reference algorithm restated in C++
    with exact 128-bit products, OpenMP over primes) on this host, at one
    core and at all cores, BSGS phase alone (primes generated beforehand) and
    end to end (prime generation + BSGS + modpow verification of every hit):
    the small base-745 workload in full (4 k, n 10000-20000, every prime of
    [1e9, 1e9+1e7]); a fixed 2^26-wide window of the R745-shaped workload
    (all cores; one core on its first primes); sr_745.abcd over the
    reference's default test range."""
    import oracle

    threads = args.cpu_threads

    def job_of(pl, mm):
        rk, rr, _ = pl.rows()
        return oracle.Job(pl.base, pl.Q, mm, pl.minN, pl.maxN, rk, rr, pl.c)

    def run(pl, mm, lo, hi, nth, limit=None):
        """BSGS-only and end-to-end seconds for the primes of [lo, hi)."""
        job = job_of(pl, mm)
        t = time.perf_counter()
        offs = rs.PrimeGen(hi, threads=nth).range(lo, hi)
        gen_s = time.perf_counter() - t
        if limit is not None:
            offs = offs[:limit]
        t = time.perf_counter()
        hits, mmc = oracle.bsgs_many(job, lo, offs, threads=nth)
        bsgs_s = time.perf_counter() - t
        t = time.perf_counter()
        for p, k, n in hits:
            oracle.verify(p, k, pl.base, n, pl.c)
        ver_s = time.perf_counter() - t
        span = (int(offs[-1]) + 1) if limit is not None else hi - lo
        if limit is not None:  # generation scaled to the primes sieved
            gen_s *= len(offs) / max(1, limit)
        e2e = gen_s + bsgs_s + ver_s
        return {"cores": nth, "primes": int(len(offs)), "p_range": span,
                "bsgs_s": bsgs_s, "end_to_end_s": e2e,
                "bsgs_p_per_sec": span / bsgs_s, "end_to_end_p_per_sec": span / e2e,
                "bsgs_primes_per_sec": len(offs) / bsgs_s, "factors": len(hits),
                "montmuls_per_prime": mmc / max(1, len(offs))}

    out = {"kind": "port", "cpu_model": cpu_model(), "threads_all": threads}
    # the R745-shaped window (the headline workload)
    lo = 10**12
    w_all = run(plan, m, lo, lo + (1 << 26), threads)
    w_one = run(plan, m, lo, lo + (1 << 26), 1, limit=60000)
    log(f"cpu: R745-shaped 2^26 window {w_all['bsgs_s']:.2f}s on {threads} threads, "
        f"{w_one['primes']} primes {w_one['bsgs_s']:.2f}s on 1")
    # the small base-745 workload in full
    c1 = rs.Plan(rs.Abcd.synth(745, 4, 3, 10**6, 10000, 20000, 1000, seed=1))
    c1m = c1.auto_m()
    c1_all = run(c1, c1m, 10**9, 10**9 + 10**7 + 1, threads)
    c1_one = run(c1, c1m, 10**9, 10**9 + 10**7 + 1, 1)
    log(f"cpu: small base-745 workload {c1_all['bsgs_s']:.2f}s / {c1_one['bsgs_s']:.2f}s")
    # sr_745.abcd over the reference's default range
    r745 = rs.Plan(rs.Abcd.read(os.path.join(ROOT, "tests", "golden", "sr_745.abcd")), 18)
    d0 = 102297149770000
    r_all = run(r745, 512, d0, d0 + 2500000, threads)
    r_one = run(r745, 512, d0, d0 + 2500000, 1)
    log(f"cpu: sr_745.abcd default range {r_all['bsgs_s']:.2f}s / {r_one['bsgs_s']:.2f}s")
    out.update({
        "value": w_all["bsgs_p_per_sec"], "unit": "p/sec", "cores": threads,
        "primes_per_sec": w_all["bsgs_primes_per_sec"],
        "sample": f"R745-shaped job (same job and m = {m}), every prime of [1e12, 1e12+2^26) "
                  f"({w_all['primes']} primes), BSGS phase, {w_all['bsgs_s']:.2f}s",
        "r745_shaped_window": {"all_cores": w_all, "one_core": w_one},
        "small_base745_full": {"workload": "base 745, 4 odd k in [3,1e6] (seed 1), n in "
                                           "[10000,20000] presieved by odd primes < 1000, every "
                                           f"prime of [1e9, 1e9+1e7], m = {c1m}",
                               "all_cores": c1_all, "one_core": c1_one},
        "sr745_default_range": {"workload": "sr_745.abcd, Q=18, m=512, every prime of "
                                            "[102297149770000, +2.5e6)",
                                "all_cores": r_all, "one_core": r_one},
        "_montmuls_per_prime": w_all["montmuls_per_prime"]})
    return out


def timed_device_chunks(rs, sv, lows, bufs):
    """Kernel-resident rate of a sieve over device-resident chunks."""
    sv.submit_device(lows[0], bufs[0])
    sv.collect()
    sv.timing()
    t0 = time.perf_counter()
    for lo, b in zip(lows, bufs):
        sv.submit_device(lo, b)
    hits = sv.collect()
    dt = time.perf_counter() - t0
    return hits, dt, sv.timing()


def bench_real_r745(args, rs, chunk):
    """The reference's own input, sr_745.abcd, at its README flags (-m 8,
    m = 512; -Q 18), from its default start p = 102297149770000, with the
    primes generated on the device beforehand."""
    path = os.path.join(ROOT, "tests", "golden", "sr_745.abcd")
    plan = rs.Plan(rs.Abcd.read(path), 18)
    p0 = 102297149770000
    K = args.real_chunks
    gpg = rs.GpuPrimeGen(p0 + K * chunk + 1)
    lows = [p0 + i * chunk for i in range(K)]
    bufs = [gpg.device_range(lo, lo + chunk) for lo in lows]
    nprimes = sum(b.count for b in bufs)
    out = {"workload": f"sr_745.abcd (the reference's test input: 22 even k, {plan.nrows} rows "
                       f"at Q=18), p in [{p0}, +{K}*2^30)", "primes": nprimes}
    m_t, t_t, _ = rs.autotune(plan, p0, bufs[0], count=min(bufs[0].count, 1 << 22))
    for name, m, tile in (("readme_m512", 512, 0), ("autotuned", m_t, t_t)):
        sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 10, hit_capacity=1 << 20)
        hits, dt, tm = timed_device_chunks(rs, sv, lows, bufs)
        out[name] = {"m": sv.m, "tile": sv.tile, "T": sv.T, "value": K * chunk / dt,
                     "unit": "p/sec", "primes_per_sec": nprimes / dt,
                     "vs_a100_published": K * chunk / dt / A100_PUBLISHED,
                     "factors": len(hits), "kernel_ms": tm["kernel_ms"]}
        if plan.contains(18300, 242670):
            raise RuntimeError("sr_745.abcd unexpectedly lists the datless debug factor")
        sv.close()
        log(f"sr_745.abcd {name}: m={out[name]['m']} tile={out[name]['tile']} "
            f"{out[name]['value']:.3e} p/s, {len(hits)} factors")
    for b in bufs:
        b.free()
    return out


def bench_edge(args, rs, chunk):
    """Many-k / large-p edge: base 2, 500 odd k in [3, 1e9], n in [1e6, 2e6]
    presieved by odd primes < 1e6, p just below 2^62 (device-generated,
    Miller-Rabin above the sieving bound).  The baby table never leaves
    LDS: the autotuner trades m against occupancy."""
    t = time.time()
    plan = rs.Plan(rs.Abcd.synth(2, 500, 3, 10**9, 1000000, 2000000, 10**6, seed=3))
    synth_s = time.time() - t
    K = args.edge_chunks
    hi = 1 << 62
    lows = [hi - (K - i) * chunk for i in range(K)]
    gpg = rs.GpuPrimeGen(hi + 1)
    t = time.time()
    bufs = [gpg.device_range(lo, lo + chunk) for lo in lows]
    gen_s = time.time() - t
    nprimes = sum(b.count for b in bufs)
    if args.edge_m:
        m, tile, cands = args.edge_m, args.edge_tile, []
    else:
        m, tile, cands = rs.autotune(plan, lows[0], bufs[0], count=min(bufs[0].count, 1 << 20),
                                     effort=1)
    sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 10, hit_capacity=1 << 20)
    hits, dt, tm = timed_device_chunks(rs, sv, lows, bufs)
    roof = roofline(rs, plan, sv, tm, K, nprimes, workload="edge")
    out = {"workload": f"base 2, 500 odd k in [3,1e9] (seed 3), n in [1e6,2e6] presieved by "
                       f"odd primes < 1e6 ({plan.nrows} rows, Q={plan.Q}), "
                       f"p in [2^62 - {K}*2^30, 2^62)",
           "value": K * chunk / dt, "unit": "p/sec", "primes_per_sec": nprimes / dt,
           "primes": nprimes, "m": sv.m, "tile": sv.tile, "T": sv.T,
           # LDS of one wavefront's layout, of one workgroup's for the
           # shared-table kernels
           "lds_bytes": sv.lds_bytes, "kernel": sv.kernel, "factors": len(hits),
           "kernel_ms": tm["kernel_ms"], "gpu_primegen_s": gen_s, "synth_s": synth_s,
           "autotune": [[cm, ct, round(ns, 3)] for cm, ct, ns in cands],
           "roofline": roof}
    log(f"edge: m={sv.m} tile={sv.tile} T={sv.T} lds={sv.lds_bytes}B "
        f"{out['value']:.3e} p/s")
    sv.close()
    for b in bufs:
        b.free()
    return out


def tune_jobs(rs, plans, p0, buf, sample):
    """Per-job (m, tile) by the autotuner's coordinate descent, and the
    measured kernel ns per prime of the choice."""
    ms, tiles, ns = [], [], []
    for pl in plans:
        m, t, cands = rs.autotune(pl, p0, buf, count=min(buf.count, sample), effort=1)
        ms.append(m)
        tiles.append(t)
        ns.append(min(c[2] for c in cands if c[0] == m and c[1] == t))
    return ms, tiles, ns


def bench_batched(args, rs, p0, chunk):
    """Heterogeneous jobs: plain back-to-back vs one graph-captured batch."""
    import workloads

    t = time.time()
    specs = workloads.job_specs(args.batch_jobs, seed=4)
    plans = [workloads.job_plan(sp) for sp in specs]
    synth_s = time.time() - t
    gpg = rs.GpuPrimeGen(p0 + chunk + 1)
    buf = gpg.device_range(p0, p0 + chunk)
    t = time.time()
    ms, tiles, _ = tune_jobs(rs, plans, p0, buf, 1 << 21)
    tune_s = time.time() - t
    log(f"batched: {len(plans)} jobs synthesized in {synth_s:.1f}s, tuned in {tune_s:.1f}s")
    svs = [rs.Sieve(pl, m=m, tile=tl, max_batch=1 << 10, hit_capacity=1 << 20)
           for pl, m, tl in zip(plans, ms, tiles)]
    plain_hits = []
    t0 = time.perf_counter()
    for sv in svs:
        sv.submit_device(p0, buf)
        plain_hits.append(sv.collect())
    plain_s = time.perf_counter() - t0
    plain_kernel_ms = sum(sv.timing()["kernel_ms"] for sv in svs)
    for sv in svs:
        sv.close()
    log(f"batched: plain back-to-back {plain_s:.3f}s")
    bat = rs.Batch(plans, ms, hit_capacity=1 << 22, tiles=tiles)
    for j in range(len(plans)):
        bat.add_chunk(j, p0, buf)
    t0 = time.perf_counter()
    bat.run()  # captures the graph (one launch per occupancy class), then launches it
    hits = bat.collect()
    batched_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    bat.run()  # replays the captured graph
    replay_hits = bat.collect()
    replay_s = time.perf_counter() - t0
    if replay_hits != hits:
        raise RuntimeError("graph replay hits differ from the first run")
    log(f"batched: one graph {batched_s:.3f}s (capture included), replay {replay_s:.3f}s")
    per_job = [[] for _ in plans]
    for (j, p, k, n) in hits:
        per_job[j].append((p, k, n))
    if per_job != plain_hits:
        raise RuntimeError("batched hits differ from the plain runs")
    bat.close()
    primes = buf.count * len(plans)
    buf.free()
    return {"workload": f"{len(plans)} heterogeneous jobs (base U[3,1030], k-count log-U[1,200], "
                        "n-width log-U[1e4,5e5] from n=1e5, presieve < 1e5; seed 4), each over "
                        "the p-chunk [1e12, 1e12+2^30)",
            "plain_s": plain_s, "batched_s": batched_s, "replay_s": replay_s,
            "plain_primes_per_sec": primes / plain_s,
            "batched_primes_per_sec": primes / batched_s,
            "replay_primes_per_sec": primes / replay_s,
            "batched_vs_plain": plain_s / batched_s, "replay_vs_plain": plain_s / replay_s,
            "plain_kernel_ms": plain_kernel_ms,
            "plain_kernel_busy_frac": plain_kernel_ms / 1e3 / plain_s,
            "factors": len(hits), "hits_equal_plain": True,
            "ms": ms, "tiles": tiles, "tune_s": tune_s}


def serve_run(rs, plans, ms, tiles, ns, p0, p_hi, chunk, arrivals, args):
    """Runs the jobs through the server; arrivals are times in seconds from
    the start.  Returns per-job latency (s), results, makespan and primes."""
    import numpy as np

    srv = rs.Server(hbm_budget=int(args.serve_budget_gib * 2**30), max_active=args.serve_active,
                    chunk_width=chunk, lds_cap=65536, hit_capacity=1 << 22,
                    quantum_ms=args.serve_quantum_ms)
    n = len(plans)
    ids, done_at = [None] * n, [None] * n
    nxt, max_active = 0, 0
    t0 = time.perf_counter()
    while True:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            ids[nxt] = srv.submit(plans[nxt], p0, p_hi, m=ms[nxt], tile=tiles[nxt],
                                  ns_per_prime=ns[nxt])
            nxt += 1
        active, queued, _ = srv.step()
        max_active = max(max_active, active)
        now = time.perf_counter() - t0
        for i in range(nxt):
            if done_at[i] is None and srv.poll(ids[i])[0] == srv.DONE:
                done_at[i] = now
        if nxt == n and all(d is not None for d in done_at):
            break
        if active == 0 and queued == 0 and nxt < n:
            time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
    results = [srv.results(j) for j in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([d - a for d, a in zip(done_at, arrivals)])
    makespan = max(done_at) - min(arrivals)
    serve_run.kernel_ms = st["kernel_ms"]  # sieve launches of this run
    return lat, results, makespan, st["primes"], max_active


def bench_serving(args, rs, p0, chunk):
    """Continuous batching vs the same jobs plain back-to-back."""
    import numpy as np

    import workloads

    specs = workloads.job_specs(args.serve_jobs, seed=6)
    plans = [workloads.job_plan(sp) for sp in specs]
    p_hi = p0 + args.serve_chunks * chunk
    gpg = rs.GpuPrimeGen(p0 + chunk + 1)
    buf = gpg.device_range(p0, p0 + chunk)
    ms, tiles, ns = tune_jobs(rs, plans, p0, buf, 1 << 21)
    buf.free()
    # plain back-to-back: one job at a time, device-generated primes per chunk
    plain_hits, plain_job_s = [], []
    plain_kernel_ms = 0.0
    t0 = time.perf_counter()
    for pl, m, tl in zip(plans, ms, tiles):
        ts = time.perf_counter()
        sv = rs.Sieve(pl, m=m, tile=tl, max_batch=1 << 10, hit_capacity=1 << 20)
        for c in range(args.serve_chunks):
            sv.submit_range(p0 + c * chunk, p0 + (c + 1) * chunk)
        plain_hits.append(sv.collect())
        tmj = sv.timing()
        primes_job = tmj["primes"]
        plain_kernel_ms += tmj["kernel_ms"]
        sv.close()
        plain_job_s.append(time.perf_counter() - ts)
    plain_s = time.perf_counter() - t0
    total_primes = primes_job * len(plans)
    log(f"serving: {len(plans)} jobs x {args.serve_chunks} chunks plain back-to-back {plain_s:.2f}s")
    out = {"workload": f"{len(plans)} jobs of the batched distribution (seed 6), each over "
                       f"[1e12, 1e12+{args.serve_chunks}*2^30) in 2^30 chunks; HBM budget "
                       f"{args.serve_budget_gib} GiB, at most {args.serve_active} active, "
                       f"windows of ~{args.serve_quantum_ms} ms device time (at most 2^30)",
           "plain_s": plain_s, "plain_primes_per_sec": total_primes / plain_s,
           "plain_sieve_kernel_busy_frac": plain_kernel_ms / 1e3 / plain_s,
           "plain_job_s_p50": float(np.percentile(plain_job_s, 50)),
           "plain_job_s_p99": float(np.percentile(plain_job_s, 99))}
    # saturated: every job arrives at t = 0
    lat, res, span, primes, act = serve_run(rs, plans, ms, tiles, ns, p0, p_hi, chunk,
                                            [0.0] * len(plans), args)
    if res != plain_hits:
        raise RuntimeError("served hits differ from the plain runs")
    if primes != total_primes:
        raise RuntimeError(f"the server counted {primes} primes, the plain runs {total_primes}")
    out["saturated"] = {"aggregate_primes_per_sec": primes / span, "makespan_s": span,
                        "latency_p50_s": float(np.percentile(lat, 50)),
                        "latency_p99_s": float(np.percentile(lat, 99)),
                        "max_active": act,
                        "sieve_kernel_busy_frac": serve_run.kernel_ms / 1e3 / span,
                        "vs_plain": plain_s / span}
    log(f"serving: saturated makespan {span:.2f}s p99 {np.percentile(lat, 99):.2f}s")
    # Poisson arrivals at a fraction of the plain job rate
    rate = args.serve_load * len(plans) / plain_s
    arr = workloads.poisson_arrivals(len(plans), rate, seed=7)
    lat, res, span, primes, act = serve_run(rs, plans, ms, tiles, ns, p0, p_hi, chunk, arr, args)
    if res != plain_hits:
        raise RuntimeError("served hits differ from the plain runs")
    out["poisson"] = {"rate_jobs_per_s": rate, "load": args.serve_load,
                      "aggregate_primes_per_sec": primes / span, "makespan_s": span,
                      "latency_p50_s": float(np.percentile(lat, 50)),
                      "latency_p99_s": float(np.percentile(lat, 99)),
                      "sieve_kernel_busy_frac": serve_run.kernel_ms / 1e3 / span,
                      "max_active": act}
    log(f"serving: poisson rate {rate:.2f}/s p50 {np.percentile(lat, 50):.2f}s "
        f"p99 {np.percentile(lat, 99):.2f}s")
    if args.serve_sweep:
        sweep = []
        for load in [float(x) for x in args.serve_sweep.split(",") if x]:
            rate_l = load * len(plans) / plain_s
            arr_l = workloads.poisson_arrivals(len(plans), rate_l, seed=7)
            lat, res, span, primes, act = serve_run(rs, plans, ms, tiles, ns, p0, p_hi, chunk,
                                                    arr_l, args)
            if res != plain_hits:
                raise RuntimeError("served hits differ from the plain runs")
            sweep.append({"load": load, "rate_jobs_per_s": rate_l,
                          "latency_p50_s": float(np.percentile(lat, 50)),
                          "latency_p99_s": float(np.percentile(lat, 99)),
                          "aggregate_primes_per_sec": primes / span, "max_active": act,
                          "vs_plain_rate": primes / span / (total_primes / plain_s)})
            log(f"serving sweep: load {load} p50 {np.percentile(lat, 50):.2f}s "
                f"p99 {np.percentile(lat, 99):.2f}s")
        out["sweep"] = sweep
    out["hits_equal_plain"] = True
    out["factors"] = sum(len(h) for h in plain_hits)
    out["ms"], out["tiles"] = ms, tiles
    return out


def bench_cli(args):
    """The reference's own flow through the drop-in host: rsieve on
    sr_745.abcd at the README's A100 flags (-b 5 -m 8 -s 192 -Q 18) from
    102297149770000 over a 2^span p-range: the CPU prime generator
    overlapped with the sieve, pinned double-buffered copies, verification
    and the factor file.  The rate is the one rsieve prints, computed as the
    reference does (kernel_v0.24.6.cu:1552-1559, :1783): p-range over the
    wall time of the loop.  Then the same with -g (primes generated on the
    GPU).  Both runs must report the same factors."""
    import re
    import subprocess
    import tempfile

    exe = os.path.join(ROOT, "cuda-riesel-sieve_amd", "rsieve")
    abcd = os.path.join(ROOT, "tests", "golden", "sr_745.abcd")
    p0 = 102297149770000
    p1 = p0 + (1 << args.cli_span_log2)
    out = {"workload": f"rsieve -i sr_745.abcd -b 5 -m 8 -s 192 -Q 18 -p {p0} -P {p1} "
                       f"(2^{args.cli_span_log2} p-range)"}
    found = {}
    for name, extra in (("cpu_primes", []), ("gpu_primes", ["-g"])):
        with tempfile.TemporaryDirectory() as d:
            ff = os.path.join(d, "factors.txt")
            cmd = [exe, "-i", abcd, "-b", "5", "-m", "8", "-s", "192", "-Q", "18", "-p", str(p0),
                   "-P", str(p1), "-O", ff] + extra
            t = time.perf_counter()
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=900)
            wall = time.perf_counter() - t
            if r.returncode != 0:
                raise RuntimeError(f"rsieve failed ({r.returncode}): {r.stderr[-500:]}")
            prog = re.search(r"Progress = (\d+) at (\d+) p/sec", r.stdout)
            secs = re.search(r"Time taken: ([0-9.e+-]+)s", r.stdout)
            nf = int(re.search(r"Factors found: (\d+)", r.stdout).group(1))
            found[name] = open(ff).read().splitlines() if os.path.exists(ff) else []
        out[name] = {"value": float(prog.group(2)), "unit": "p/sec",
                     "p_range": int(prog.group(1)), "loop_s": float(secs.group(1)),
                     "process_wall_s": wall, "factors": nf,
                     "vs_a100_published": float(prog.group(2)) / A100_PUBLISHED,
                     "flags": " ".join(a for a in cmd[1:] if a not in ("-O", ff))}
        log(f"cli {name}: {out[name]['value']:.3e} p/s, {nf} factors, {wall:.1f}s")
    if sorted(found["cpu_primes"]) != sorted(found["gpu_primes"]):
        raise RuntimeError("rsieve with CPU and GPU primes wrote different factors")
    return out


def bench_msweep(args, rs, p0):
    """Table sizes m = 2^10 .. 2^14 for the R745-shaped job on the primes of
    [1e12, 1e12+2^26): one wavefront per prime with its own table (tiles 1,
    2, 4) and one table shared by the 4 or 8 wavefronts of a workgroup
    (tiles 1, 2); kernel ns per prime, or why a layout is refused."""
    plan = rs.Plan(rs.Abcd.synth(745, 22, 3, 10**7, 180000, 250000, 10**5, seed=args.seed))
    gpg = rs.GpuPrimeGen(p0 + (1 << 26) + 1)
    buf = gpg.device_range(p0, p0 + (1 << 26))
    out = []
    G4, G8 = rs.TILE_GROUP4, rs.TILE_GROUP8
    for m in (1024, 2048, 4096, 8192, 16384):
        for tile in (1, 2, 4, 1 | G4, 2 | G4, 1 | G8, 2 | G8):
            rec = {"m": m, "tile": tile}
            try:
                sv = rs.Sieve(plan, m=m, tile=tile, max_batch=1 << 10, hit_capacity=1 << 20)
            except rs.RsError as e:
                rec["refused"] = rs.strerror(e.status)
                out.append(rec)
                continue
            sv.submit_device(p0, buf)  # warm-up
            sv.collect()
            sv.timing()
            sv.submit_device(p0, buf)
            sv.collect()
            tm = sv.timing()
            rec.update({"T": sv.T, "kernel": sv.kernel, "lds_bytes": sv.lds_bytes,
                        "ns_per_prime": tm["kernel_ms"] * 1e6 / buf.count})
            sv.close()
            out.append(rec)
            log(f"msweep: m={m} tile={tile} T={rec['T']} {rec['ns_per_prime']:.3f} ns/prime "
                f"({rec['kernel']}, {rec['lds_bytes']} B)")
    buf.free()
    return out


if __name__ == "__main__":
    main()
