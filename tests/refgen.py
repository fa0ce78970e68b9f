"""Test inputs reproduced bit for bit from the reference's own test suites.

The reference draws its random test panels and reads from an in-test
xorshift64* generator with fixed seeds (crates/eschalign/src/simd/mod.rs:685-738,
crates/eschalign/tests/gpu_parity.rs:34-123). Porting the generator and the
helpers exactly gives this suite the very inputs those tests check.
"""
from __future__ import annotations

import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASK = (1 << 64) - 1
ACGT = b"ACGT"


class Rng:
    """xorshift64* (simd/mod.rs:685-699, gpu_parity.rs:34-48)."""

    def __init__(self, seed: int):
        self.s = seed & MASK

    def next(self) -> int:
        x = self.s
        x ^= x >> 12
        x ^= (x << 25) & MASK
        x ^= x >> 27
        self.s = x
        return (x * 0x2545F4914F6CDD1D) & MASK

    def below(self, n: int) -> int:
        return self.next() % n

    def chance(self, p: float) -> bool:
        return (self.next() >> 11) / float(1 << 53) <= p


def random_seq(rng: Rng, length: int, n_rate: float) -> bytes:
    """simd/mod.rs:702-712 / gpu_parity.rs:50-60."""
    out = bytearray()
    for _ in range(length):
        if rng.chance(n_rate):
            out.append(ord("N"))
        else:
            out.append(ACGT[rng.below(4)])
    return bytes(out)


def mutate(rng: Rng, s: bytes) -> bytes:
    """simd/mod.rs:715-731 / gpu_parity.rs:63-79: ~4% del, 5% sub, 3% ins."""
    out = bytearray()
    for b in s:
        if rng.chance(0.04):
            continue
        if rng.chance(0.05):
            out.append(ACGT[rng.below(4)])
        else:
            out.append(b)
        if rng.chance(0.03):
            out.append(ACGT[rng.below(4)])
    return bytes(out)


def random_panel_simd(rng: Rng, n_refs: int):
    """simd/mod.rs:733-738: lengths 20 + below(280), 1% N."""
    refs = []
    for _ in range(n_refs):
        length = 20 + rng.below(280)
        refs.append(random_seq(rng, length, 0.01))
    return refs


def random_panel(rng: Rng, n_refs: int, min_len: int, spread: int):
    """gpu_parity.rs:81-87."""
    refs = []
    for _ in range(n_refs):
        length = min_len + rng.below(spread)
        refs.append(random_seq(rng, length, 0.01))
    return refs


def reads_for(rng: Rng, panel, n: int):
    """gpu_parity.rs:106-126: every length 1..=40, then n reads drawn from
    the panel (3 of 4) or random (1 of 4). Returns letters."""
    reads = [random_seq(rng, length, 0.02) for length in range(1, 41)]
    for t in range(n):
        if t % 4 == 3:
            length = 1 + rng.below(300)
            reads.append(random_seq(rng, length, 0.02))
        else:
            r = panel[rng.below(len(panel))]
            a = rng.below(len(r) // 3 + 1)
            b = len(r) - rng.below(len(r) // 3 + 1)
            lead, tail = rng.below(30), rng.below(30)
            read = random_seq(rng, lead, 0.0)
            read += mutate(rng, r[a:max(b, a)])
            read += random_seq(rng, tail, 0.0)
            reads.append(read)
    return reads


def simd_random_reads(rng: Rng, panel, count: int = 40):
    """simd/mod.rs:761-773 (the read mix of simd_matches_scalar_random)."""
    reads = []
    for t in range(count):
        if t % 4 == 3:
            length = 1 + rng.below(300)
            reads.append(random_seq(rng, length, 0.02))
        else:
            r = panel[rng.below(len(panel))]
            a = rng.below(len(r) // 3 + 1)
            b = len(r) - rng.below(len(r) // 3 + 1)
            lead, tail = rng.below(30), rng.below(30)
            read = random_seq(rng, lead, 0.0)
            read += mutate(rng, r[a:max(b, a)])
            read += random_seq(rng, tail, 0.0)
            reads.append(read)
    return reads


def chimeric_read(rng: Rng, panel, length: int) -> bytes:
    """simd/mod.rs:928-938."""
    read = bytearray()
    while len(read) < length:
        gap = rng.below(400)
        read += random_seq(rng, gap, 0.01)
        r = panel[rng.below(len(panel))]
        read += mutate(rng, r)
    return bytes(read[:length])


def read_fasta(path):
    """(name, letters) in file order, the rule of lib/fasta.rs:42-93."""
    recs = []
    cur = None
    with open(path, "rb") as fh:
        for line in fh:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                if cur is not None:
                    recs.append(cur)
                name = line[1:].split()[0].decode()
                cur = (name, bytearray())
            elif cur is not None:
                cur[1].extend(b for b in line if not chr(b).isspace())
    if cur is not None:
        recs.append(cur)
    return [(n, bytes(s)) for n, s in recs]


def read_fastq(path):
    recs = []
    with open(path, "rb") as fh:
        lines = [ln.rstrip(b"\r\n") for ln in fh]
    for k in range(0, len(lines) - 3, 4):
        recs.append((lines[k][1:].decode(), lines[k + 1]))
    return recs


def golden():
    with open(os.path.join(GOLDEN, "align_golden.json")) as fh:
        return json.load(fh)


def fixture_reads():
    """The golden's non-random queries: the real tRNA reads
    (gpu_parity.rs:190-205)."""
    g = golden()
    return [p["query"].encode() for p in g["pairs"] if p["source"] != "random"]


def fixture_panel(name: str):
    return [s for _, s in read_fasta(os.path.join(GOLDEN, name))]


SCHEMES_SIMD = [(2, -1, -10, -1), (1, -1, -2, -1), (3, -2, -5, -2)]
SCHEMES_GPU = SCHEMES_SIMD + [(3, 1, -5, -2), (2, -1, 0, 0)]
SCHEMES_TRANSPOSED = SCHEMES_SIMD + [(2, 0, -10, -1), (3, 1, -5, -2)]


def encode(letters: bytes) -> np.ndarray:
    """lib/alphabet.rs:29-55 restated with a lookup table (host helper)."""
    t = np.full(256, 4, dtype=np.uint8)
    for ch, c in ((b"A", 0), (b"C", 1), (b"G", 2), (b"T", 3), (b"U", 3)):
        t[ch[0]] = c
        t[ch.lower()[0]] = c
    return t[np.frombuffer(letters, dtype=np.uint8)] if len(letters) else np.zeros(0, np.uint8)


def revcomp_codes(codes: np.ndarray) -> np.ndarray:
    c = codes[::-1].copy()
    m = c != 4
    c[m] = 3 - c[m]
    return c
