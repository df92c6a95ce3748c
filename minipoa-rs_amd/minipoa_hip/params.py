"""Params / Sequence / AlignedBase mirroring minipoa-rs src/lib.rs:39-145."""
from __future__ import annotations

from collections import namedtuple
from dataclasses import dataclass, field, replace

# src/lib.rs:140-145 -- None encodes Rust's Option::None.
AlignedBase = namedtuple("AlignedBase", "node base aligned")


def default_matrix():
    """src/lib.rs:73-80: ACGT match 2 / mismatch -4, N row/column 0, N-N 2."""
    m = [0] * 25
    for i in range(4):
        for j in range(4):
            m[i * 5 + j] = 2 if i == j else -4
    m[24] = 2
    return m


@dataclass
class Params:
    """src/lib.rs:47-107 Params with the reference's defaults."""

    match_score: int = 2
    mismatch: int = -4
    gap_open: int = -4
    gap_ext: int = -2
    b: int = 100
    f: int = 40
    adaptive: bool = False
    seeding: bool = False
    k: int = 19
    w: int = 10
    progressive: bool = False
    threads: int = 1
    win_band: int = 500
    win_max: int = 10000
    mm_filter_ratio: float = 0.25
    max_occ: int = 0
    tr_dist: int = 1000
    result: int = 0
    verbose: int = 0
    alphabet: bytes = b"ACGTN"
    matrix: list = field(default_factory=default_matrix)
    is_rna: bool = False
    # Cargo feature `scalar-dp` (Cargo.toml:19-25): force the i64 kernel.
    scalar_dp: bool = False

    def clone(self, **kw):
        c = replace(self, **kw)
        if "matrix" not in kw:
            c.matrix = list(self.matrix)
        return c

    # src/lib.rs:110-126
    def encode(self, base):
        b = base if isinstance(base, int) else ord(base)
        if 97 <= b <= 122:
            b -= 32
        if len(self.alphabet) <= 5 and b == ord("U"):
            b = ord("T")
        a = bytes(self.alphabet)
        if b in a:
            return a.index(b)
        if ord("N") in a:
            return a.index(ord("N"))
        return len(a) - 1

    # src/lib.rs:127-134
    def decode(self, base):
        c = self.alphabet[base]
        if self.is_rna and len(self.alphabet) <= 5 and c == ord("T"):
            return ord("U")
        return c

    # src/lib.rs:135-137
    def score(self, a, b):
        return self.matrix[a * len(self.alphabet) + b]

    def rebuild_matrix(self):
        """src/cli.rs:188-199: diagonal = match, last row/column = 0, else mismatch."""
        n = len(self.alphabet)
        self.matrix = [
            self.match_score if i == j else (0 if i == n - 1 or j == n - 1 else self.mismatch)
            for i in range(n) for j in range(n)
        ]
        return self


@dataclass
class Sequence:
    """src/lib.rs:39-45"""

    name: str = ""
    comment: str = ""
    seq: bytes = b""
    qual: bytes = b""


def records(seqs, prefix="read"):
    return [Sequence(name=f"{prefix}{i}", seq=bytes(s)) for i, s in enumerate(seqs)]
