"""Launch-shape autotuner for the fill kernel.

For each row-width bucket it runs a calibration batch whose DP rows fall in
the bucket, once per candidate launch shape (waves per job x columns per
lane) and LDS ring size, and keeps the fastest by wall time of the whole
batch (the engine, the fill and the traceback together: the quantity that
matters). The table it writes is read by the library when MPOA_TUNING names
the file; every shape computes the same DP (the parity tests cover each), so
tuning changes speed only.

    python -m minipoa_hip.tune tuning.txt          # about three minutes
    MPOA_TUNING=tuning.txt python bench.py

File format: one line per bucket, "max_width cols waves ring_kb", in
increasing max_width; lines starting with # are comments.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

from . import Context, Params, batch_run, synth

SHAPES = [(8, 1), (16, 1), (4, 2), (8, 2), (4, 4), (8, 4)]
RINGS_KB = [16, 24, 32]
BUCKETS = [384, 768, 1024, 2048, 4096]


def calibration(width, cells=4e9, reads=6, seed=7):
    """Adaptive-band DNA jobs whose rows are about `width` columns wide:
    template length 2 x width (at least 600), band b chosen so that
    2 (b + m/40) + 16 is about `width`. Returns (params, job_off, read_off,
    bases)."""
    L = max(600, 2 * width)
    b = max(8, width // 2 - L // 40 - 8)
    per_job = (reads - 1) * 1.3 * L * width
    njobs = int(min(1024, max(16, cells / per_job)))
    jo, ro, bs = synth.jobs(seed=seed, njobs=njobs, lmin=L, lmax=L, rmin=reads, rmax=reads,
                            emin=0.10, emax=0.10)
    return Params(adaptive=True, b=b), jo, ro, bs


def measure(ctx, params, jo, ro, bs, shape, ring_kb, reps=2):
    """Best wall seconds of `reps` runs of the batch with a forced shape and ring."""
    saved = {k: os.environ.get(k) for k in ("MPOA_FILL_SHAPE", "MPOA_RING_KB", "MPOA_TUNING")}
    os.environ["MPOA_FILL_SHAPE"] = f"{shape[0]},{shape[1]}"
    os.environ["MPOA_RING_KB"] = str(ring_kb)
    os.environ.pop("MPOA_TUNING", None)
    try:
        best, cells = None, 0
        for _ in range(reps):
            t = time.perf_counter()
            r = batch_run(ctx, params, jo, ro, bs)
            dt = time.perf_counter() - t
            cells = r.stats()["cells"]
            del r
            best = dt if best is None else min(best, dt)
        return best, cells
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def tune(ctx, path, buckets=BUCKETS, shapes=SHAPES, rings=RINGS_KB, cells=4e9, reps=2, log=print):
    """Measures every (shape, ring) per bucket and writes the table to `path`.
    Returns the chosen [(max_width, cols, waves, ring_kb, gcups)]."""
    chosen = []
    for width in buckets:
        params, jo, ro, bs = calibration(width, cells=cells)
        batch_run(ctx, params, jo, ro, bs)  # warm buffers for this size
        best = None
        for shape in shapes:
            for ring in rings:
                dt, n = measure(ctx, params, jo, ro, bs, shape, ring, reps)
                g = n / dt / 1e9
                log(f"width <= {width}: shape {shape[0]}x{shape[1]} ring {ring} KB: "
                    f"{dt * 1e3:.0f} ms, {g:.1f} GCUPS")
                if best is None or dt < best[0]:
                    best = (dt, shape, ring, g)
        chosen.append((width, best[1][0], best[1][1], best[2], round(best[3], 2)))
    with open(path, "w") as f:
        f.write("# max_width cols waves ring_kb   (minipoa_hip.tune; GCUPS of the calibration batch)\n")
        for w, c, v, r, g in chosen:
            f.write(f"{w} {c} {v} {r}   # {g} GCUPS\n")
    return chosen


def main(argv):
    if len(argv) != 2:
        print("usage: python -m minipoa_hip.tune OUT_FILE", file=sys.stderr)
        return 2
    ctx = Context(0)
    try:
        for row in tune(ctx, argv[1], log=lambda s: print(s, file=sys.stderr, flush=True)):
            print(*row)
    finally:
        ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
