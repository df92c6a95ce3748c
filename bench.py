#!/usr/bin/env python3
"""Throughput of batched progressive POA on one MI355X.

Workload: 4,096 independent DNA POA jobs, template length U[200, 2000],
U[8, 32] reads per job at U[5%, 15%] error (60/20/20 sub/ins/del), default
scores with the adaptive band, consensus + MSA output. One step = the whole
workload (every progressive-insertion step of every job). By default the
step runs device-resident (mpoa_batch_run_resident): one launch of the
persistent kernel, in which each workgroup takes jobs from a device queue and
runs each to its end -- graph fusion, re-ranking and DP planning, fill,
traceback, once per read. `--engine host` runs the host-driven engine
(mpoa_batch_run) instead: fill and traceback launches on several streams,
planning and fusion on host threads. Inputs are generated on the host before
timing; the timed region covers staging, every DP, fusion, and rendering.

Prints one JSON line. value = aggregate GCUPS (DP cells of the reference's
exact band geometry / wall second).

Besides the headline workload, `legs` reports the other workloads once each,
every one checked byte for byte against the oracle:
  c1  plain unseeded consensus: one 2,000 bp template, 10 reads at 10%,
      one DP per launch (the drop-in path), beside the oracle on one core;
  c3  seeded long-read consensus (-S -r2): one 16,500 bp template, 100 reads
      at 10% trimmed by up to 500 bases at each end; each read aligns as
      ~28 minimizer-anchored windows that share one launch per step; then
      16 such jobs in one batch beside the oracle on all host cores;
  c4  protein POA: 2,000 families over a 20-letter alphabet with a drawn
      symmetric matrix, 16 members at 30% divergence, full DP (-f0);
  c5  serving: Poisson arrivals of DNA jobs (template log-uniform in
      [300, 20000] bp, 5-64 reads at 5-15% error; 10% extend a 50-read graph
      loaded from GFA text (-i) by 20 reads), admitted by device footprint;
      p50/p99 job latency;
  rec the headline workload once more with the DP in the records layout
      (1-8 byte traceback records instead of M/D/I values): its step time,
      traceback bytes per cell, and its outputs against the values run;
  host the headline workload on the host-driven engine (when the headline
      is device-resident), on a context of its own: step time and outputs;
  res the headline workload device-resident (when the headline is the host
      engine).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

# At least eight hardware queues per process (HIP opens four by default),
# set before HIP starts: the engine then runs five launch slots, each stream
# on a queue of its own. A larger setting is kept. `--hw-queues N` sets N
# exactly (4: the library default a drop-in caller gets, three slots).
try:
    _queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "4"))
except ValueError:
    _queues = 4
_queues = max(_queues, 8)
for _k, _a in enumerate(sys.argv):
    if _a == "--hw-queues" and _k + 1 < len(sys.argv):
        _queues = int(sys.argv[_k + 1])
    elif _a.startswith("--hw-queues="):
        _queues = int(_a.split("=", 1)[1])
os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(_queues, 1), 32))

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "minipoa-rs_amd"), str(ROOT / "oracle")]

METRIC = ("aggregate GCUPS + POA jobs/sec (batched vs plain, bit-exact); "
          "p50/p99 job latency")
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 12  # M, D and I (i32) stored per cell for the traceback (values layout)
VALU_PEAK_TOPS = 256 * 4 * 16 * 2.4e9 / 1e12  # int32 lane-ops/s, nominal


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """Host cores this process may use: the affinity mask, bounded by the
    cgroup CPU quota when one is set (a quota of q cores time-slices more
    threads onto q cores). Returns (cores, description)."""
    affinity = len(os.sched_getaffinity(0))
    quota = None
    try:
        q, period = Path("/sys/fs/cgroup/cpu.max").read_text().split()[:2]
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    cores = affinity if quota is None else max(1, min(affinity, int(quota)))
    model = ""
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return cores, {"os_cpu_count": os.cpu_count(), "affinity": affinity,
                   "cgroup_quota_cores": quota, "model": model}


def int_valu_peak():
    """Measured int32 VALU rate (lane-ops/s) of device 0: dependency-free
    add/max chains (minipoa-rs_amd/csrc/probe_valu.hip), best of 5 runs of
    2,048 workgroups x 256 lanes x 8 chains x 32,768 steps."""
    import ctypes

    lib = ctypes.CDLL(str(ROOT / "minipoa-rs_amd" / "libmpoa_probe.so"))
    ops, ms = ctypes.c_double(0), ctypes.c_double(0)
    rc = lib.mpoa_probe_int_valu(0, 2048, 32768, 5, ctypes.byref(ops), ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"int-VALU probe failed (HIP error {rc})")
    log(f"int-VALU probe: {ops.value/1e12:.2f} T lane-ops/s ({ms.value:.2f} ms per run)")
    return ops.value


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
    ap.add_argument("--engine", choices=["resident", "host"], default="resident",
                    help="headline path: device-resident (mpoa_batch_run_resident, default) or the "
                         "host-driven engine (mpoa_batch_run)")
    ap.add_argument("--resident", dest="engine", action="store_const", const="resident",
                    help="same as --engine resident")
    ap.add_argument("--hw-queues", type=int, default=None,
                    help="hardware queues per process (read before HIP starts; default 8)")
    ap.add_argument("--no-valu-probe", dest="valu_probe", action="store_false",
                    help="skip the int-VALU rate probe (the roofline then uses the nominal peak)")
    ap.add_argument("--legs", default="rec,other,c1,c3,c4,c5",
                    help="other workloads to report once each (rec, host, res, other = the path the "
                         "headline does not take, c1, c3, c4, c5; empty = none)")
    ap.add_argument("--serve-jobs", type=int, default=96)
    ap.add_argument("--serve-rate", type=float, default=6.0, help="serving arrivals per second")
    ap.add_argument("--serve-slots", type=int, default=0, help="serving launch slots (0 = library default)")
    args = ap.parse_args()
    args.resident = args.engine == "resident"
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
    valu_peak = int_valu_peak() if args.valu_probe else None

    for w in range(args.warmup):
        t = time.time()
        r = mh.batch_run(ctx, params, job_off, read_off, bases, resident=args.resident)
        log(f"warmup {w}: {time.time() - t:.2f}s")
        del r

    times, stats, res = [], [], None
    for s in range(args.steps):
        res = None  # free the previous step's graphs outside the timed region
        t = time.perf_counter()
        res = mh.batch_run(ctx, params, job_off, read_off, bases, resident=args.resident)
        dt = time.perf_counter() - t
        times.append(dt)
        stats.append(res.stats())
        if args.resident:
            log(f"step {s}: {dt:.3f}s cells={stats[-1]['cells']/1e9:.2f}G device-resident: "
                f"device={stats[-1]['resident_device_ms']:.0f}ms resident={stats[-1]['resident_jobs']} "
                f"evicted={stats[-1]['evicted_jobs']} launches={stats[-1]['step_replays']}")
            continue
        log(f"step {s}: {dt:.3f}s cells={stats[-1]['cells']/1e9:.2f}G "
            f"fill={stats[-1]['fill_ms']:.0f}ms device={stats[-1]['device_ms']:.0f}ms "
            f"host={stats[-1]['host_ms']:.0f}ms (plan {stats[-1]['plan_ms']:.0f} "
            f"consume {stats[-1]['consume_ms']:.0f} fuse {stats[-1]['fuse_ms']:.0f} "
            f"render {stats[-1]['render_ms']:.0f} pack/slot {stats[-1]['pack_ms']:.0f}) "
            f"launches={stats[-1]['fill_launches']}")
    st = stats[-1]
    failed = sum(res.status(j) != 0 for j in range(args.jobs))
    total = sum(times)
    cells = st["cells"]
    gcups = cells * len(times) / total / 1e9
    jobs_per_s = args.jobs * len(times) / total

    # Dominant kernel: the fill. Algorithmic bytes: the M, D and I scores
    # (4 B each) written per cell for the traceback. Predecessor rows are
    # re-read from the LDS ring, so they add no HBM bytes by design.
    # Algorithmic int ops per cell: 6 (profile add, md max, I add/add/max,
    # M max) + 5 per predecessor overlap (M max, D add/add/max/max) + 2 for
    # the adaptive argmax.
    fill_s = st["fill_ms"] / 1e3
    launches = st["fill_launches"]
    wall_s = total / len(times)
    # The fills report the traceback bytes they wrote (12 per i32 cell in
    # the values layout, which the headline runs in).
    alg_bytes = st["arena_bytes"] or BYTES_PER_CELL * cells
    int_ops = 6 * cells + 5 * st["pred_cells"] + 2 * cells
    # The device-resident headline is one launch of the persistent kernel
    # per step (step, fill and traceback of every job inside it), bracketed
    # by HIP events on its own stream: that launch is the dominant kernel. A
    # run of the captured step loops (MPOA_RES_PERSIST=0) has its fills as
    # nodes of replayed HIP graphs, which events cannot bracket one by one:
    # only the wall-based figures are reported for it.
    persistent = args.resident and st["buckets"] == 1 and st["step_replays"] == 1
    kernel = "poa_fill"
    if persistent:
        kernel = "rj_persist_kernel"
        fill_s = st["resident_device_ms"] / 1e3
        launches = 1
    per_launch = fill_s > 0
    achieved_gbs = alg_bytes / fill_s / 1e9 if per_launch else None
    rnd = lambda x, n: None if x is None else round(x, n)
    traffic = None
    tfile = ROOT / "profiles" / ("rj_persist_traffic.json" if persistent else "poa_fill_traffic.json")
    if tfile.exists() and (persistent or not args.resident):
        # HBM bytes per cell measured by rocprofv3 FETCH_SIZE and WRITE_SIZE
        # passes over this workload (copy-kernel scaled), times this run's
        # cells per launch: profile-derived, not measured in this run.
        t = json.loads(tfile.read_text())
        traffic = int((t["fetch_bytes_per_cell"] + t["write_bytes_per_cell"]) * cells
                      / max(1, launches))
    peak_t = (valu_peak or VALU_PEAK_TOPS * 1e12) / 1e12
    roofline = {
        "bound": "hbm", "achieved": rnd(achieved_gbs, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": rnd(achieved_gbs and achieved_gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
        "traffic_source": f"profile-derived (profiles/{tfile.name} x this run's cells)",
        "kernel": kernel, "launches": launches,
        "avg_launch_ms": round(fill_s * 1e3 / max(1, launches), 3) if per_launch else None,
        "bytes_per_launch": alg_bytes // max(1, launches),
        "bytes_per_cell": round(alg_bytes / max(1, cells), 3),
        # Over the wall: the step's algorithmic bytes per wall second (the
        # per-launch figure divides one launch's bytes by its own time, and
        # launches of the engine's streams overlap).
        "wall_gbs": round(alg_bytes / wall_s / 1e9, 1),
        "wall_frac": round(alg_bytes / wall_s / 1e9 / HBM_PEAK_GBS, 4),
        # The DP is integer max-plus: int VALU is the binding roof.
        "valu": {"achieved_tops": rnd(int_ops / fill_s / 1e12 if per_launch else None, 3),
                 "peak_tops": round(peak_t, 2),
                 "peak_source": "measured on this device (add/max probe)" if valu_peak
                 else "nominal 256 CU x 64 lanes x 2.4 GHz",
                 "peak_nominal_tops": round(VALU_PEAK_TOPS, 1),
                 "frac": rnd(int_ops / fill_s / 1e12 / peak_t if per_launch else None, 4),
                 "wall_tops": round(int_ops / wall_s / 1e12, 3),
                 "wall_frac": round(int_ops / wall_s / 1e12 / peak_t, 4),
                 "ops_per_cell": round(int_ops / max(1, cells), 2)},
    }

    out = {
        "metric": METRIC, "value": round(gcups, 3), "unit": "GCUPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total / len(times) * 1e3, 1), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i32", "data": "synthetic",
        "config": {"workload": "batched DNA POA, adaptive band: 4096 jobs, L U[200,2000], "
                               "8-32 reads/job, 5-15% error 60/20/20, consensus+MSA",
                   "jobs": args.jobs, "reads": int(job_off[-1]), "cells": cells,
                   "mode": "device-resident (mpoa_batch_run_resident)" if args.resident
                   else "host-driven engine (mpoa_batch_run)"},
        "jobs_per_s": round(jobs_per_s, 1), "failed_jobs": failed,
        "dp_calls": st["dp_calls"], "retries": st["retries"], "i64_calls": st["i64_calls"],
        "kernel_gcups": round(cells / fill_s / 1e9, 2) if per_launch else None,
        "host_ms": round(st["host_ms"], 1), "device_ms": round(st["device_ms"], 1),
        "roofline": roofline,
        "hip_hw_queues": int(os.environ["GPU_MAX_HW_QUEUES"]),
    }

    if args.cpu_jobs > 0:
        import oracle

        n = min(args.cpu_jobs, args.jobs)
        cores, host = host_cores()
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
            "kind": "port", "host": host, "jobs_per_s": round(n / dt, 1),
            "sample": f"first {n} of the {args.jobs} jobs ({ccells/1e9:.2f} G cells), "
                      f"oracle 8-lane i32 restatement, {dt:.1f}s"}
        out["bit_exact"] = {"jobs_checked": n, "identical": exact,
                            "what": "FNV-1a of consensus FASTA + MSA per job, device vs oracle"}
        out["vs_cpu"] = round(gcups / (ccells / dt / 1e9), 1)
        log(f"cpu baseline: {n} jobs {dt:.1f}s, bit-exact {exact}/{n}")
    legs = [x for x in args.legs.split(",") if x]
    if legs:
        import oracle

        cores, _ = host_cores()
        out["legs"] = {}
        if "c4" in legs:
            out["legs"]["c4"] = leg_c4(ctx, oracle, cores)
        if "c5" in legs:
            out["legs"]["c5"] = leg_c5(ctx, oracle, cores, args.serve_jobs, args.serve_rate,
                                       args.serve_slots)
        if "c3" in legs:
            out["legs"]["c3"] = leg_c3(ctx, oracle, cores)
        other = "host" if args.resident else "res"
        for name in ("host", "res"):
            if name in legs or ("other" in legs and name == other):
                out["legs"][name] = leg_path(name == "res", params, job_off, read_off, bases, res, gcups,
                                             total / len(times))
        # One DP per launch on the context's own stream last: after that
        # stream has run, the batched legs on the same context measured
        # 40% slower (C4 690 ms instead of 400-450 ms).
        if "c1" in legs:
            out["legs"]["c1"] = leg_c1(ctx, oracle)
        # On a context of its own, after the legs on the main context (its
        # streams would share the hardware queues with theirs).
        if "rec" in legs:
            out["legs"]["rec"] = leg_records(params, job_off, read_off, bases, res, gcups, total / len(times))
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
    ctx.close()
    print(json.dumps(out), flush=True)


def fnv1a_one(text):
    """FNV-1a 64 of one long text (a plain loop: megabytes in well under a second)."""
    h = 14695981039346656037
    for c in text.encode():
        h = ((h ^ c) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def hashes_of(get_text, n):
    import numpy as np

    return np.concatenate([fnv1a_many([get_text(j).encode() for j in range(c, min(n, c + 512))])
                           for c in range(0, n, 512)]) if n else np.zeros(0, dtype=np.uint64)


def leg_records(params, job_off, read_off, bases, values_res, values_gcups, values_s):
    """The headline workload on the host-driven engine in the records layout
    (one warmup, one timed run) on a context of its own: step time against
    the headline's, traceback bytes per cell, and every job's consensus + MSA
    against the headline run's."""
    import numpy as np

    import minipoa_hip as mh

    c = mh.Context(0)
    c.set_layout(mh.LAYOUT_RECORDS)
    mh.batch_run(c, params, job_off, read_off, bases)
    t = time.perf_counter()
    r = mh.batch_run(c, params, job_off, read_off, bases)
    dt = time.perf_counter() - t
    st = r.stats()
    n = len(job_off) - 1
    want = hashes_of(lambda j: values_res.output(j, 0) + values_res.output(j, 1), n)
    got = hashes_of(lambda j: r.output(j, 0) + r.output(j, 1), n)
    same = int((got == want).sum())
    out = {"mode": "host-driven engine, records layout", "ms_per_step": round(dt * 1e3, 1),
           "gcups": round(st["cells"] / dt / 1e9, 3),
           "vs_headline": round(st["cells"] / dt / 1e9 / values_gcups, 3),
           "headline_ms_per_step": round(values_s * 1e3, 1),
           "fill_ms": round(st["fill_ms"], 1), "record_launches": st["record_launches"],
           "fill_launches": st["fill_launches"],
           "bytes_per_cell": round(st["arena_bytes"] / max(1, st["cells"]), 3),
           "identical_to_headline": same, "jobs": n}
    del r
    c.close()
    log(f"rec: records layout {dt*1e3:.0f} ms ({out['vs_headline']:.2f}x headline), "
        f"{out['bytes_per_cell']} B/cell, identical {same}/{n}")
    return out


def leg_path(resident, params, job_off, read_off, bases, head_res, head_gcups, head_s):
    """The headline workload on the path the headline does not take (one
    warmup, two timed runs, on a context of its own): device-resident
    (mpoa_batch_run_resident) or the host-driven engine (mpoa_batch_run).
    Step time against the headline's, and every job's consensus + MSA
    against the headline run's."""
    import minipoa_hip as mh

    # A context of its own, closed afterwards: its device buffers (arenas,
    # graph blocks) do not stay beside the main context's.
    ctx = mh.Context(0)
    mh.batch_run(ctx, params, job_off, read_off, bases, resident=resident)
    times = []
    for _ in range(2):
        t = time.perf_counter()
        r = mh.batch_run(ctx, params, job_off, read_off, bases, resident=resident)
        times.append(time.perf_counter() - t)
        st = r.stats()
        if len(times) < 2:
            del r
    dt = min(times)
    n = len(job_off) - 1
    want = hashes_of(lambda j: head_res.output(j, 0) + head_res.output(j, 1), n)
    got = hashes_of(lambda j: r.output(j, 0) + r.output(j, 1), n)
    same = int((got == want).sum())
    out = {"mode": "device-resident (mpoa_batch_run_resident)" if resident
           else "host-driven engine (mpoa_batch_run)",
           "ms_per_step": round(dt * 1e3, 1), "ms_runs": [round(x * 1e3, 1) for x in times],
           "gcups": round(st["cells"] / dt / 1e9, 3),
           "vs_headline": round(st["cells"] / dt / 1e9 / head_gcups, 3),
           "headline_ms_per_step": round(head_s * 1e3, 1),
           "host_ms": round(st["host_ms"], 1),
           "identical_to_headline": same, "jobs": n}
    if resident:
        out.update({"resident_jobs": st["resident_jobs"], "evicted_jobs": st["evicted_jobs"],
                    "buckets": st["buckets"], "step_replays": st["step_replays"],
                    "resident_device_ms": round(st["resident_device_ms"], 1)})
    else:
        out.update({"fill_ms": round(st["fill_ms"], 1), "fill_launches": st["fill_launches"],
                    "device_ms": round(st["device_ms"], 1)})
    del r
    ctx.close()
    log(f"{'res' if resident else 'host'}: {out['mode']} {dt*1e3:.0f} ms ({out['vs_headline']:.2f}x headline), "
        f"identical {same}/{n}")
    return out


def leg_c1(ctx, oracle):
    """Plain unseeded consensus (-r0): one job, one DP per launch."""
    import minipoa_hip as mh
    from minipoa_hip import Params, records, synth

    reads = synth.c1()
    seqs = records(reads)
    p = Params()
    mh.align_sequences(ctx, p, seqs)  # warm
    t = time.perf_counter()
    g = mh.align_sequences(ctx, p, seqs)
    dt = time.perf_counter() - t
    oracle.reset_cells()
    t = time.perf_counter()
    og = oracle.align_sequences(p, seqs)
    ct = time.perf_counter() - t
    cells = oracle.cells()
    same = g.consensus(p) == og.consensus(p) and g.msa(p, seqs) == og.msa(p, seqs)
    log(f"c1: gpu {dt*1e3:.1f} ms, oracle {ct*1e3:.1f} ms, {cells/1e6:.1f} M cells, identical {same}")
    return {"workload": "plain unseeded consensus: 2000 bp template, 10 reads at 10% error, "
                        "one DP per launch",
            "cells": cells, "gpu_ms": round(dt * 1e3, 2), "gpu_gcups": round(cells / dt / 1e9, 4),
            "cpu_ms": round(ct * 1e3, 2), "cpu_gcups": round(cells / ct / 1e9, 4), "cpu_cores": 1,
            "cpu_kind": "port", "identical": bool(same)}


def leg_c3(ctx, oracle, cores):
    """Seeded long-read consensus (-S -r2): one job through the engine; every
    read's anchored windows run side by side in one launch per step."""
    import numpy as np

    import minipoa_hip as mh
    from minipoa_hip import Params, synth

    jo, ro, b = synth.seeded()
    p = Params(seeding=True)
    r = mh.batch_run(ctx, p, jo, ro, b)  # warm
    del r
    t = time.perf_counter()
    r = mh.batch_run(ctx, p, jo, ro, b)
    dt = time.perf_counter() - t
    st = r.stats()
    text = r.output(0, 0) + r.output(0, 1)
    gfa = r.output(0, 2)
    t = time.perf_counter()
    cells, want, want_gfa = oracle.batch_run(p, jo, ro, b, 1, gfa=True)
    ct = time.perf_counter() - t
    # The reference's -t N: the windows of each read on N threads
    # (src/parallel.rs:46-67), results in window order.
    pt = Params(seeding=True, threads=cores)
    t = time.perf_counter()
    cells_t, want_t = oracle.batch_run(pt, jo, ro, b, 1)
    ctt = time.perf_counter() - t
    same = fnv1a_one(text) == int(want[0]) == int(want_t[0]) and st["cells"] == cells == cells_t
    same_gfa = fnv1a_one(gfa) == int(want_gfa[0])
    same = same and same_gfa
    nreads = int(jo[1] - jo[0])
    log(f"c3: gpu {dt*1e3:.0f} ms {st['cells']/dt/1e9:.2f} GCUPS, {st['dp_calls']} DPs; "
        f"oracle {ct:.1f}s on 1 core, {ctt:.1f}s with -t {cores}; identical {same} (GFA {same_gfa})")
    # Many such jobs at once: every ready job's windows share launches.
    nb = 16
    bj, br, bb = synth.seeded(seed=20264, njobs=nb)
    rb = mh.batch_run(ctx, p, bj, br, bb)  # warm
    del rb
    t = time.perf_counter()
    rb = mh.batch_run(ctx, p, bj, br, bb)
    bdt = time.perf_counter() - t
    bst = rb.stats()
    got = np.array([fnv1a_one(rb.output(k, 0) + rb.output(k, 1)) for k in range(nb)], dtype=np.uint64)
    del rb
    t = time.perf_counter()
    bcells, bwant = oracle.batch_run(p, bj, br, bb, cores)
    bct = time.perf_counter() - t
    bsame = int((got == bwant).sum())
    log(f"c3 batch: {nb} jobs gpu {bdt*1e3:.0f} ms {bst['cells']/bdt/1e9:.2f} GCUPS; oracle {bct:.1f}s "
        f"on {cores} cores ({bcells/bct/1e9:.2f} GCUPS); bit-exact {bsame}/{nb}")
    return {"workload": "seeded long-read consensus: 16500 bp template, 100 reads at 10% error, "
                        "ends trimmed U[0,500], minimizer windows (k19 w10), -r2",
            "cells": st["cells"], "dp_calls": st["dp_calls"], "windows_per_read": round(st["dp_calls"] / (nreads - 1), 1),
            "gpu_ms": round(dt * 1e3, 1), "gpu_gcups": round(st["cells"] / dt / 1e9, 3),
            "ms_per_read": round(dt * 1e3 / nreads, 2), "gfa_bytes": len(gfa),
            "gfa_identical": bool(same_gfa),
            "cpu_ms": round(ct * 1e3, 1), "cpu_gcups": round(cells / ct / 1e9, 3), "cpu_cores": 1,
            "cpu_threads_ms": round(ctt * 1e3, 1), "cpu_threads": cores,
            "cpu_kind": "port", "identical": bool(same),
            "batch": {"jobs": nb, "cells": bst["cells"], "gpu_ms": round(bdt * 1e3, 1),
                      "gpu_gcups": round(bst["cells"] / bdt / 1e9, 3),
                      "cpu_ms": round(bct * 1e3, 1), "cpu_gcups": round(bcells / bct / 1e9, 3),
                      "cpu_cores": cores, "bit_exact": bsame,
                      "what": "16 such jobs (seed 20264) in one batch; oracle job-parallel on host cores"}}


def leg_c4(ctx, oracle, cores):
    """Protein families, full DP, custom matrix (-m file), batched."""
    import minipoa_hip as mh
    from minipoa_hip import Params, synth

    p, jo, ro, b = synth.protein_families()
    # The drawn matrix goes through the -m file format and io::load_matrix.
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        f = Path(d) / "protein.mtx"
        f.write_text(synth.matrix_text(p.alphabet, p.matrix))
        p = mh.load_matrix(Params(f=0), f)
    n = len(jo) - 1
    r = mh.batch_run(ctx, p, jo, ro, b)  # warm
    del r
    t = time.perf_counter()
    r = mh.batch_run(ctx, p, jo, ro, b)
    dt = time.perf_counter() - t
    st = r.stats()
    t = time.perf_counter()
    ccells, want = oracle.batch_run(p, jo, ro, b, cores)
    ct = time.perf_counter() - t
    got = hashes_of(lambda j: r.output(j, 0) + r.output(j, 1), n)
    exact = int((got == want).sum())
    log(f"c4: {n} families {dt*1e3:.0f} ms {st['cells']/dt/1e9:.2f} GCUPS; oracle {ct:.1f}s on {cores} cores; "
        f"bit-exact {exact}/{n}")
    return {"workload": "protein POA, full DP (-f0): 2000 families, root U[100,800], 16 members at 30% "
                        "divergence, 20-letter alphabet, drawn symmetric matrix",
            "cells": st["cells"], "ms": round(dt * 1e3, 1), "gcups": round(st["cells"] / dt / 1e9, 3),
            "jobs_per_s": round(n / dt, 1), "fill_launches": st["fill_launches"],
            "cpu_gcups": round(ccells / ct / 1e9, 3), "cpu_jobs_per_s": round(n / ct, 1),
            "cpu_cores": cores, "cpu_kind": "port", "bit_exact": {"jobs": n, "identical": exact}}


def leg_c5(ctx, oracle, cores, njobs, rate, slots=0):
    """Serving: jobs submitted at their Poisson arrival times."""
    import numpy as np

    import minipoa_hip as mh
    from minipoa_hip import Params, Sequence, records, synth

    p = Params(adaptive=True)
    stream = synth.serving_stream(njobs=njobs, rate_per_s=rate)
    # The incremental jobs' first 50 reads exist as GFA files before the
    # stream (-r2 output of an earlier run); each job loads its graph with
    # read_gfa (-i), outside the measured stream.
    bases_of, recs_of = {}, {}
    with mh.Server(ctx, p, render=False) as pre:
        tk = {k: pre.submit(j["base"]) for k, j in enumerate(stream) if j["base"]}
        for k, t in tk.items():
            pre.wait(t)
            text = pre.graph(t).gfa(p, records(stream[k]["base"]))
            bases_of[k], recs_of[k] = mh.read_gfa(p, text)
    def run_stream(**kw):
        srv = mh.Server(ctx, p, slots=slots, **kw)
        tickets = []
        t0 = time.perf_counter()
        for k, j in enumerate(stream):
            delay = j["t"] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            g = bases_of.get(k)
            tickets.append(srv.submit(j["reads"], graph=g, existing=len(recs_of.get(k, []))))
        for t in tickets:
            srv.wait(t)
        wall = time.perf_counter() - t0
        times = [srv.times(t) for t in tickets]
        lat = np.array([x["done_ms"] - x["submit_ms"] for x in times])
        qwait = np.array([x["admit_ms"] - x["submit_ms"] for x in times])
        first = min(x["submit_ms"] for x in times)
        span = (max(x["done_ms"] for x in times) - first) / 1e3
        cells = int(sum(x["cells"] for x in times))
        failed = sum(srv.status(t) != 0 for t in tickets)
        got = hashes_of(lambda k: srv.output(tickets[k], 0) + srv.output(tickets[k], 1), len(tickets))
        st = srv.stats()
        srv.close()
        return lat, qwait, span, cells, failed, got, st, wall

    lat, qwait, span, cells, failed, got, st, wall = run_stream()
    rlat, rqwait, rspan, rcells, rfailed, rgot, rst, rwall = run_stream(resident=True)
    # Oracle: ordinary jobs from scratch (job-parallel); incremental jobs
    # continue a field-by-field copy of the graph read_gfa loaded.
    plain = [k for k, j in enumerate(stream) if not j["base"]]
    full = [stream[k]["reads"] for k in plain]
    jo = np.zeros(len(full) + 1, dtype=np.uint64)
    jo[1:] = np.cumsum([len(r) for r in full])
    lens = [len(x) for r in full for x in r]
    ro = np.zeros(len(lens) + 1, dtype=np.uint64)
    ro[1:] = np.cumsum(lens)
    bb = np.frombuffer(b"".join(x for r in full for x in r), dtype=np.uint8)

    def extended(k):
        g, rec = bases_of[k], recs_of[k]
        nodes, aligned, order = g.snapshot()
        og = oracle.graph_from_arrays(p, nodes, aligned, order, g.paths)
        seqs = rec + [
            Sequence(name=f"read{len(rec) + t}", seq=r) for t, r in enumerate(stream[k]["reads"])]
        oracle.extend_graph(p, og, seqs, len(rec))
        return fnv1a_one(">consensus sequence\n" + og.consensus(p)[0].decode() + "\n" + og.msa(p, seqs))

    from concurrent.futures import ThreadPoolExecutor

    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        inc = {k: ex.submit(extended, k) for k in bases_of}
        _, want_plain = oracle.batch_run(p, jo, ro, bb, cores)
        want = np.zeros(len(stream), dtype=np.uint64)
        for i, k in enumerate(plain):
            want[k] = want_plain[i]
        for k, f in inc.items():
            want[k] = f.result()
    ct = time.perf_counter() - t
    exact = int((got == want).sum())
    rexact = int((rgot == want).sum())
    pct = lambda a, q: round(float(np.percentile(a, q)), 1)
    log(f"c5 resident: wall {rwall:.1f}s, p50 {pct(rlat, 50)} ms p99 {pct(rlat, 99)} ms, "
        f"{rcells/rspan/1e9:.2f} GCUPS, bit-exact {rexact}/{len(stream)}, resident {rst['resident_jobs']} "
        f"evicted {rst['evicted_jobs']}")
    log(f"c5: {len(stream)} jobs over {stream[-1]['t']:.1f}s of arrivals, wall {wall:.1f}s, "
        f"p50 {pct(lat, 50)} ms p99 {pct(lat, 99)} ms, {cells/span/1e9:.2f} GCUPS, bit-exact {exact}/{len(stream)} "
        f"(oracle {ct:.1f}s)")
    return {"workload": f"serving: Poisson arrivals at {rate}/s, {len(stream)} DNA jobs, template log-uniform "
                        "[300,20000] bp, 5-64 reads at 5-15% error, 10% extend a 50-read graph loaded from GFA "
                        "(-i) by 20 reads, adaptive band",
            "p50_ms": pct(lat, 50), "p99_ms": pct(lat, 99), "max_ms": round(float(lat.max()), 1),
            "admission_wait_p99_ms": pct(qwait, 99), "jobs_per_s": round(len(stream) / span, 2),
            "gcups": round(cells / span / 1e9, 3), "cells": cells, "span_s": round(span, 2),
            "fill_launches": st["fill_launches"], "failed_jobs": int(failed),
            "incremental_jobs": len(bases_of),
            "bit_exact": {"jobs": len(stream), "identical": exact,
                          "what": "FNV-1a of consensus FASTA + MSA per job, served vs oracle (from scratch, "
                                  "or continuing the GFA-loaded graph)"},
            # The same stream served on device-resident graphs (size-class step
            # tables, captured step graphs; mpoa_server_opts.resident).
            "resident": {"p50_ms": pct(rlat, 50), "p99_ms": pct(rlat, 99), "max_ms": round(float(rlat.max()), 1),
                         "admission_wait_p99_ms": pct(rqwait, 99), "jobs_per_s": round(len(stream) / rspan, 2),
                         "gcups": round(rcells / rspan / 1e9, 3), "span_s": round(rspan, 2),
                         "resident_jobs": rst["resident_jobs"], "evicted_jobs": rst["evicted_jobs"],
                         "step_replays": rst["step_replays"], "failed_jobs": int(rfailed),
                         "bit_exact": {"jobs": len(stream), "identical": rexact}}}


if __name__ == "__main__":
    main()
