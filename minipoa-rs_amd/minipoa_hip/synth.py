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


def seeded(seed=20263, njobs=1, length=16500, nreads=100, rate=0.10, trim=500):
    """Seeded long-read consensus: per job one template of `length` bases and
    `nreads` reads at `rate`, each trimmed by U[0, trim] bases at both ends.
    Returns (job_off, read_off, bases)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    per_job = []
    for _ in range(njobs):
        tmpl = rng.integers(0, 4, size=length).astype(np.uint8)
        reads = []
        for _ in range(nreads):
            r = mutate(rng, tmpl, rate)
            a = int(rng.integers(0, trim + 1))
            b = int(rng.integers(0, trim + 1))
            reads.append(r[a:max(a, len(r) - b)])
        per_job.append(reads)
    return _pack(per_job)


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


def _pack(job_reads_list):
    """List of per-job read lists -> (job_off, read_off, bases)."""
    chunks, lens, nreads = [], [], []
    for reads in job_reads_list:
        nreads.append(len(reads))
        for r in reads:
            chunks.append(r)
            lens.append(len(r))
    jo = np.zeros(len(nreads) + 1, dtype=np.uint64)
    jo[1:] = np.cumsum(nreads)
    ro = np.zeros(len(lens) + 1, dtype=np.uint64)
    ro[1:] = np.cumsum(lens)
    return jo, ro, np.frombuffer(b"".join(chunks) or b"\0", dtype=np.uint8).copy()


PROTEIN = b"ARNDCQEGHILKMFPSTWYV"


def matrix_text(alphabet, matrix):
    """A scoring matrix in the -m file format (header of labels, then one
    labelled row per letter)."""
    n = len(alphabet)
    lines = ["# drawn by minipoa_hip.synth", " ".join(chr(c) for c in alphabet)]
    for i in range(n):
        lines.append(chr(alphabet[i]) + " " + " ".join(str(matrix[i * n + j]) for j in range(n)))
    return "\n".join(lines) + "\n"


def protein_matrix(rng, alphabet=PROTEIN, dlo=4, dhi=11, olo=-4, ohi=3):
    """Symmetric integer matrix: diagonal U[dlo,dhi], off-diagonal U[olo,ohi]."""
    n = len(alphabet)
    m = np.zeros((n, n), dtype=np.int64)
    for i in range(n):
        m[i, i] = rng.integers(dlo, dhi + 1)
        for j in range(i + 1, n):
            m[i, j] = m[j, i] = rng.integers(olo, ohi + 1)
    return [int(x) for x in m.reshape(-1)]


def diverge(rng, root, nsym, rate, sub=0.6, ins=0.2, indel_mean=3.0):
    """root: int array of symbol codes. round(rate*L) events at distinct
    positions: substitutions (sub, uniform over the other symbols), insertions
    of a Geometric(indel_mean) run of uniform symbols before the position
    (ins), or deletions of the Geometric(indel_mean) run starting at the
    position (the rest; runs are clipped at the end and may cover other
    events' positions, whose insertions are kept)."""
    L = len(root)
    k = min(int(round(rate * L)), L)
    t = np.asarray(root, dtype=np.int64).copy()
    if k == 0:
        return t
    pos = np.sort(rng.choice(L, size=k, replace=False))
    u = rng.random(k)
    is_sub = u < sub
    is_ins = (u >= sub) & (u < sub + ins)
    is_del = u >= sub + ins
    sp = pos[is_sub]
    t[sp] = (t[sp] + rng.integers(1, nsym, size=len(sp))) % nsym
    keep = np.ones(L, dtype=bool)
    dp = pos[is_del]
    dl = rng.geometric(1.0 / indel_mean, size=len(dp))
    if len(dp):
        starts = np.repeat(dp, dl)
        within = np.arange(int(dl.sum())) - np.repeat(np.cumsum(dl) - dl, dl)
        idx = starts + within
        keep[idx[idx < L]] = False
    ip = pos[is_ins]
    il = rng.geometric(1.0 / indel_mean, size=len(ip))
    syms = rng.integers(0, nsym, size=int(il.sum()))
    at = np.repeat(ip, il)
    t = np.insert(t, at, syms)
    keep = np.insert(keep, at, True)
    return t[keep]


def protein_families(seed=20264, nfam=2000, lmin=100, lmax=800, members=16, rate=0.30):
    """Protein POA with a custom matrix and full DP (-f0): per family a root
    of length U[lmin,lmax] over a 20-letter alphabet and `members` sequences
    at `rate` divergence (60% substitutions, 20% insertions, 20% deletions,
    indel lengths Geometric with mean 3). Returns (params, job_off,
    read_off, bases); params has f = 0 and the drawn matrix."""
    from .params import Params

    rng = np.random.Generator(np.random.PCG64(seed))
    alpha = np.frombuffer(PROTEIN, dtype=np.uint8)
    matrix = protein_matrix(rng)
    fams = []
    for _ in range(nfam):
        L = int(rng.integers(lmin, lmax + 1))
        root = rng.integers(0, len(PROTEIN), size=L)
        fams.append([alpha[diverge(rng, root, len(PROTEIN), rate)].tobytes() for _ in range(members)])
    params = Params(f=0, alphabet=PROTEIN, matrix=matrix)
    jo, ro, b = _pack(fams)
    return params, jo, ro, b


def serving_stream(seed=20265, njobs=96, rate_per_s=6.0, lmin=300, lmax=20000, rmin=5, rmax=64,
                   emin=0.05, emax=0.15, incr_frac=0.10, incr_base=50, incr_add=20):
    """Serving workload: Poisson arrivals at rate_per_s; template length
    log-uniform in [lmin, lmax]; U[rmin, rmax] reads at U[emin, emax] error
    (60/20/20). A fraction incr_frac of the jobs are incremental: a graph of
    incr_base reads exists before the stream (as loaded with -i) and the job
    adds incr_add more reads. Returns a list of dicts with keys t (arrival, s),
    reads (the job's reads), base (reads already in the graph, [] if none)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t = 0.0
    jobs = []
    for _ in range(njobs):
        t += float(rng.exponential(1.0 / rate_per_s))
        L = int(round(np.exp(rng.uniform(np.log(lmin), np.log(lmax)))))
        incr = bool(rng.random() < incr_frac)
        R = incr_base + incr_add if incr else int(rng.integers(rmin, rmax + 1))
        tmpl = rng.integers(0, 4, size=L).astype(np.uint8)
        reads = [mutate(rng, tmpl, float(rng.uniform(emin, emax))) for _ in range(R)]
        if incr:
            jobs.append({"t": t, "base": reads[:incr_base], "reads": reads[incr_base:]})
        else:
            jobs.append({"t": t, "base": [], "reads": reads})
    return jobs
