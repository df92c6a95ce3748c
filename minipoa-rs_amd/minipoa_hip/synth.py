"""Seeded synthetic POA workloads (DNA reads drawn from random templates).

A read of template length L at error rate r gets round(r*L) error events at
distinct uniform template positions; each event is a substitution (60%,
uniform over the three other bases), an insertion before the position (20%,
uniform ACGT) or a deletion (20%). All randomness comes from numpy's PCG64
with fixed seeds, so every workload is reproducible byte for byte.
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def mutate(rng, template_codes, rate, sub=0.6, ins=0.2):
    """template_codes: uint8 array of 0..3. Returns uint8 ASCII read."""
    L = len(template_codes)
    k = int(round(rate * L))
    t = template_codes.copy()
    if k:
        pos = np.sort(rng.choice(L, size=min(k, L), replace=False))
        kind = rng.random(len(pos))
        is_sub = kind < sub
        is_ins = (kind >= sub) & (kind < sub + ins)
        is_del = kind >= sub + ins
        sp = pos[is_sub]
        t[sp] = (t[sp] + rng.integers(1, 4, size=len(sp))) % 4
        ip = pos[is_ins]
        ib = rng.integers(0, 4, size=len(ip)).astype(np.uint8)
        keep = np.ones(L, dtype=bool)
        keep[pos[is_del]] = False
        # insertions go before their position; a deleted position keeps its
        # insertion slot (there is none: events are distinct per position).
        t = np.insert(t, ip, ib)
        keep = np.insert(keep, ip, True)
        t = t[keep]
    return ACGT[t].tobytes()


def c1(seed=20261, length=2000, nreads=10, rate=0.10):
    """Plain unseeded consensus: one 2,000 bp template, 10 reads at 10%."""
    rng = np.random.Generator(np.random.PCG64(seed))
    tmpl = rng.integers(0, 4, size=length).astype(np.uint8)
    return [mutate(rng, tmpl, rate) for _ in range(nreads)]


def jobs(seed=20262, njobs=4096, lmin=200, lmax=2000, rmin=8, rmax=32, emin=0.05, emax=0.15):
    """Batched DNA POA: per job a template length U[lmin,lmax], U[rmin,rmax]
    reads, per-read error rate U[emin,emax]. Returns (job_off, read_off,
    bases) as numpy arrays (offsets into reads / into bases)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    chunks, read_lens, job_reads = [], [], []
    for _ in range(njobs):
        L = int(rng.integers(lmin, lmax + 1))
        R = int(rng.integers(rmin, rmax + 1))
        tmpl = rng.integers(0, 4, size=L).astype(np.uint8)
        for _ in range(R):
            rate = float(rng.uniform(emin, emax))
            r = mutate(rng, tmpl, rate)
            chunks.append(r)
            read_lens.append(len(r))
        job_reads.append(R)
    job_off = np.zeros(njobs + 1, dtype=np.uint64)
    job_off[1:] = np.cumsum(job_reads)
    read_off = np.zeros(len(read_lens) + 1, dtype=np.uint64)
    read_off[1:] = np.cumsum(read_lens)
    bases = np.frombuffer(b"".join(chunks), dtype=np.uint8).copy()
    return job_off, read_off, bases


def subset(job_off, read_off, bases, idx):
    """The jobs `idx` of a workload as a new workload."""
    chunks, lens, nreads = [], [], []
    for j in idx:
        r0, r1 = int(job_off[j]), int(job_off[j + 1])
        nreads.append(r1 - r0)
        for r in range(r0, r1):
            a, b = int(read_off[r]), int(read_off[r + 1])
            chunks.append(bases[a:b].tobytes())
            lens.append(b - a)
    jo = np.zeros(len(nreads) + 1, dtype=np.uint64)
    jo[1:] = np.cumsum(nreads)
    ro = np.zeros(len(lens) + 1, dtype=np.uint64)
    ro[1:] = np.cumsum(lens)
    return jo, ro, np.frombuffer(b"".join(chunks), dtype=np.uint8).copy()


def job_reads(job_off, read_off, bases, j):
    r0, r1 = int(job_off[j]), int(job_off[j + 1])
    return [bases[int(read_off[r]):int(read_off[r + 1])].tobytes() for r in range(r0, r1)]
