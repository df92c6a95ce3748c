#!/usr/bin/env python3
"""Throughput of batched progressive POA on one MI355X.

Workload: 4,096 independent DNA POA jobs, template length U[200, 2000],
U[8, 32] reads per job at U[5%, 15%] error (60/20/20 sub/ins/del), default
scores with the adaptive band, consensus + MSA output. One step = the whole
workload (every progressive-insertion step of every job: one fill launch and
one traceback launch per step across all jobs). Inputs are generated on the
host before timing; the timed region covers planning, device DP, host graph
fusion and rendering.

Prints one JSON line. value = aggregate GCUPS (DP cells of the reference's
exact band geometry / wall second).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "minipoa-rs_amd"), str(ROOT / "oracle")]

METRIC = ("aggregate GCUPS + POA jobs/sec (batched vs plain, bit-exact); "
          "p50/p99 job latency")
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # int32 lane-ops/s, nominal


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fnv1a_many(texts):
    """FNV-1a 64 of many byte strings at once (numpy, lockstep over jobs)."""
    import numpy as np

    n = len(texts)
    lens = np.array([len(t) for t in texts], dtype=np.int64)
    width = int(lens.max()) if n else 0
    buf = np.zeros((n, width), dtype=np.uint8)
    for k, t in enumerate(texts):
        buf[k, :len(t)] = np.frombuffer(t, dtype=np.uint8)
    h = np.full(n, 14695981039346656037, dtype=np.uint64)
    prime = np.uint64(1099511628211)
    with np.errstate(over="ignore"):
        for c in range(width):
            live = lens > c
            nh = (h ^ buf[:, c].astype(np.uint64)) * prime
            h = np.where(live, nh, h)
    return h


def copy_calibration(mib):
    """A device-to-device copy of `mib` MiB: a plain copy kernel of known size
    whose rocprofv3 FETCH_SIZE / WRITE_SIZE scale the fill kernel's counters
    to bytes."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (a, b):
        if hip.hipMalloc(ctypes.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed for the copy calibration")
    hip.hipMemset(a, 1, n)
    hip.hipDeviceSynchronize()
    rc = hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    if rc != 0:
        raise RuntimeError("hipMemcpy failed for the copy calibration")
    log(f"copy calibration: {mib} MiB device-to-device")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--jobs", type=int, default=4096)
    ap.add_argument("--cpu-jobs", type=int, default=4096,
                    help="jobs of the workload timed on host cores (0 = skip)")
    ap.add_argument("--plain-jobs", type=int, default=4,
                    help="jobs run one DP per launch for the plain-path rate (0 = skip)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a device-to-device copy of this size first (counter scaling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(jobs would shard by HBM footprint with no collectives).", file=sys.stderr)
        sys.exit(2)

    import numpy as np

    import minipoa_hip as mh
    from minipoa_hip import Params, records, synth

    t0 = time.time()
    job_off, read_off, bases = synth.jobs(seed=20262, njobs=args.jobs)
    log(f"workload: {args.jobs} jobs, {int(job_off[-1])} reads, {len(bases)} bases "
        f"({time.time() - t0:.1f}s to generate)")
    params = Params(adaptive=True)
    ctx = mh.Context(0)
    if args.calib_copy_mib > 0:
        copy_calibration(args.calib_copy_mib)

    for w in range(args.warmup):
        t = time.time()
        r = mh.batch_run(ctx, params, job_off, read_off, bases)
        log(f"warmup {w}: {time.time() - t:.2f}s")
        del r

    times, stats, res = [], [], None
    for s in range(args.steps):
        res = None  # free the previous step's graphs outside the timed region
        t = time.perf_counter()
        res = mh.batch_run(ctx, params, job_off, read_off, bases)
        dt = time.perf_counter() - t
        times.append(dt)
        stats.append(res.stats())
        log(f"step {s}: {dt:.3f}s cells={stats[-1]['cells']/1e9:.2f}G "
            f"fill={stats[-1]['fill_ms']:.0f}ms device={stats[-1]['device_ms']:.0f}ms "
            f"host={stats[-1]['host_ms']:.0f}ms launches={stats[-1]['fill_launches']}")
    st = stats[-1]
    failed = sum(res.status(j) != 0 for j in range(args.jobs))
    total = sum(times)
    cells = st["cells"]
    gcups = cells * len(times) / total / 1e9
    jobs_per_s = args.jobs * len(times) / total

    # Dominant kernel: the fill. Algorithmic bytes: the M, D and I scores
    # (4 B each) written per cell for the traceback. Predecessor rows are
    # re-read from the LDS ring, so they add no HBM bytes by design.
    fill_s = st["fill_ms"] / 1e3
    launches = st["fill_launches"]
    alg_bytes = 12 * cells
    int_ops = 6 * cells + 5 * st["pred_cells"] + 2 * cells
    achieved_gbs = alg_bytes / fill_s / 1e9
    traffic = None
    tfile = ROOT / "profiles" / "poa_fill_traffic.json"
    if tfile.exists():
        # HBM bytes per cell measured by rocprofv3 FETCH_SIZE and WRITE_SIZE
        # passes over this workload (copy-kernel scaled), times this run's
        # cells per launch.
        t = json.loads(tfile.read_text())
        traffic = int((t["fetch_bytes_per_cell"] + t["write_bytes_per_cell"]) * cells
                      / max(1, launches))
    roofline = {
        "bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
        "kernel": "poa_fill", "launches": launches,
        "avg_launch_ms": round(st["fill_ms"] / max(1, launches), 3),
        "bytes_per_launch": alg_bytes // max(1, launches),
        "valu": {"achieved_tops": round(int_ops / fill_s / 1e12, 3),
                 "peak_tops": round(VALU_PEAK_TOPS, 1),
                 "frac": round(int_ops / fill_s / 1e12 / VALU_PEAK_TOPS, 4)},
    }

    out = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total / len(times) * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i32", "data": "synthetic",
        "config": {"workload": "batched DNA POA, adaptive band: 4096 jobs, L U[200,2000], "
                               "8-32 reads/job, 5-15% error 60/20/20, consensus+MSA",
                   "jobs": args.jobs, "reads": int(job_off[-1]), "cells": cells},
        "jobs_per_s": round(jobs_per_s, 1), "failed_jobs": failed,
        "dp_calls": st["dp_calls"], "retries": st["retries"], "i64_calls": st["i64_calls"],
        "fill_gcups": round(cells / fill_s / 1e9, 2),
        "host_ms": round(st["host_ms"], 1), "device_ms": round(st["device_ms"], 1),
        "roofline": roofline,
    }

    if args.plain_jobs > 0:
        # Plain path: the same jobs, one DP per launch (mpoa_align_sequences).
        n = args.plain_jobs
        sj, sr, sb = synth.subset(job_off, read_off, bases, range(n))
        t = time.perf_counter()
        for j in range(n):
            mh.align_sequences(ctx, params, records(synth.job_reads(sj, sr, sb, j)))
        dt = time.perf_counter() - t
        batch = mh.batch_run(ctx, params, sj, sr, sb)
        pcells = batch.stats()["cells"]
        out["plain"] = {"jobs": n, "gcups": round(pcells / dt / 1e9, 4),
                        "jobs_per_s": round(n / dt, 3)}
        out["batched_over_plain"] = round(gcups / (pcells / dt / 1e9), 1)
        log(f"plain: {n} jobs {dt:.2f}s")

    if args.cpu_jobs > 0:
        import oracle

        n = min(args.cpu_jobs, args.jobs)
        cores = len(os.sched_getaffinity(0))
        cores = min(cores, 16)
        sj, sr, sb = synth.subset(job_off, read_off, bases, range(n))
        t = time.perf_counter()
        ccells, hashes = oracle.batch_run(params, sj, sr, sb, cores)
        dt = time.perf_counter() - t
        got = np.concatenate([
            fnv1a_many([(res.output(j, 0) + res.output(j, 1)).encode()
                        for j in range(c, min(n, c + 512))])
            for c in range(0, n, 512)])
        exact = int((got == hashes).sum())
        out["cpu_baseline"] = {
            "value": round(ccells / dt / 1e9, 4), "unit": "GCUPS", "cores": cores,
            "kind": "port",
            "sample": f"first {n} of the {args.jobs} jobs ({ccells/1e9:.2f} G cells), "
                      f"oracle 8-lane i32 restatement, {dt:.1f}s"}
        out["bit_exact"] = {"jobs_checked": n, "identical": exact,
                            "what": "FNV-1a of consensus FASTA + MSA per job, device vs oracle"}
        out["vs_cpu"] = round(gcups / (ccells / dt / 1e9), 1)
        log(f"cpu baseline: {n} jobs {dt:.1f}s, bit-exact {exact}/{n}")
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
