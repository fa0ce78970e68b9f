"""The unpaired run from BAM input to BAM output records, around the
device mapping: the reference's CLI pipeline for one batch
(crates/eschalign-cli/src/align.rs: read_batches 604-672, process_chunk
1504-1602, build_records 1084-1196, the coordinate sort 2418-2600), with
the scoring, selection and tracebacks of esch_map_reads and the record
building and layout of esch_bam_build_records.

BGZF input is read with gzip (a BGZF file is a series of gzip members);
records stay in the BAM byte layout throughout. The output record stream
is what the CLI writes after its header, uncompressed, so its FNV-1a
digest is comparable with the reference's own digests of its output
(crates/eschalign-cli/tests/align_unpaired.rs:40-63).
"""
import ctypes as C
import gzip
import struct
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import HIT_DTYPE, READ_MAPPING_DTYPE, Panel, _check, encode, load_library

SEQ_LETTERS = b"=ACMGRSVTWYHKDBN"
FLAG_UNMAPPED, FLAG_REVERSE, FLAG_SECONDARY, FLAG_SUPPLEMENTARY = 0x4, 0x10, 0x100, 0x800


def read_fasta(path: str) -> List[Tuple[str, bytes]]:
    """lib/fasta.rs:42-104: (name, letters) per record, in file order; the
    name is the header's first word, sequence lines are concatenated with
    whitespace dropped, case kept."""
    records: List[Tuple[str, bytes]] = []
    cur = None
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        for line in f:
            line = line.rstrip(b"\r\n")
            if line.startswith(b">"):
                if cur is not None:
                    records.append((cur[0], bytes(cur[1])))
                words = line[1:].split()
                if not words:
                    raise ValueError("a FASTA header with an empty name")
                cur = (words[0].decode(), bytearray())
            elif cur is not None:
                cur[1].extend(b for b in line if b not in b" \t\r\n\v\f")
            elif line.strip():
                raise ValueError("FASTA content before the first header")
    if cur is not None:
        records.append((cur[0], bytes(cur[1])))
    return records


@dataclass
class BamFile:
    header: bytes            # magic through the reference list
    ref_names: List[str]
    records: List[bytes]     # each record without its block_size


def read_bam(path: str) -> BamFile:
    raw = gzip.open(path, "rb").read()
    if raw[:4] != b"BAM\x01":
        raise ValueError(f"{path} is not a BAM file")
    off = 8 + struct.unpack_from("<I", raw, 4)[0]
    n_ref = struct.unpack_from("<I", raw, off)[0]
    off += 4
    names = []
    for _ in range(n_ref):
        ln = struct.unpack_from("<I", raw, off)[0]
        names.append(raw[off + 4:off + 4 + ln - 1].decode())
        off += 4 + ln + 4
    header = raw[:off]
    recs = []
    while off < len(raw):
        bs = struct.unpack_from("<I", raw, off)[0]
        recs.append(raw[off + 4:off + 4 + bs])
        off += 4 + bs
    return BamFile(header, names, recs)


def record_flag(rec: bytes) -> int:
    return struct.unpack_from("<H", rec, 14)[0]


def record_name(rec: bytes) -> bytes:
    return rec[32:32 + rec[8] - 1]


def record_letters(rec: bytes) -> bytes:
    """SEQ as letters (=ACMGRSVTWYHKDBN)."""
    l_name, n_cigar, l_seq = rec[8], struct.unpack_from("<H", rec, 12)[0], struct.unpack_from("<I", rec, 16)[0]
    p = 32 + l_name + 4 * n_cigar
    packed = np.frombuffer(rec, np.uint8, (l_seq + 1) // 2, p)
    codes = np.empty(2 * len(packed), np.uint8)
    codes[0::2] = packed >> 4
    codes[1::2] = packed & 0xF
    return np.frombuffer(SEQ_LETTERS, np.uint8)[codes[:l_seq]].tobytes()


def primary_records(records: Sequence[bytes]) -> List[bytes]:
    """The reader's filter (align.rs:624-641): secondary and supplementary
    records are skipped; a record aligned to the reverse strand is refused."""
    out = []
    for r in records:
        f = record_flag(r)
        if f & (FLAG_SECONDARY | FLAG_SUPPLEMENTARY):
            continue
        if not f & FLAG_UNMAPPED and f & FLAG_REVERSE:
            raise ValueError(f"input record {record_name(r).decode(errors='replace')} is aligned to the "
                             "reverse strand; unaligned or forward-strand records only")
        out.append(r)
    return out


def build_records(records: Sequence[bytes], maps: np.ndarray, hits: np.ndarray, ops: np.ndarray,
                  panel: Panel, secondary: bool = False) -> Tuple[bytes, int]:
    """esch_bam_build_records: the output record stream (each record with
    its block_size) and the record count. maps / hits / ops are esch_mapped's
    arrays (READ_MAPPING_DTYPE, HIT_DTYPE, u32), one mapping per record."""
    L = load_library()
    n = len(records)
    stream = b"".join(struct.pack("<I", len(r)) + r for r in records)
    maps = np.ascontiguousarray(maps, READ_MAPPING_DTYPE)
    hits = np.ascontiguousarray(hits, HIT_DTYPE) if len(hits) else np.zeros(1, HIT_DTYPE)
    ops = np.ascontiguousarray(ops, np.uint32) if len(ops) else np.zeros(1, np.uint32)
    assert len(maps) == n
    letters = [np.frombuffer(s, np.uint8) for s in panel.seqs]
    lptr = (C.c_void_p * len(letters))(*[x.ctypes.data for x in letters])
    lens = (C.c_size_t * len(letters))(*[len(x) for x in letters])
    names = (C.c_char_p * len(letters))(*[s.encode() for s in panel.names])
    out_len, n_out = C.c_size_t(), C.c_size_t()
    cap = 2 * len(stream) + 4096
    while True:
        buf = np.zeros(cap, np.uint8)
        st = L.esch_bam_build_records(stream, len(stream), n, maps.ctypes.data if n else None, hits.ctypes.data,
                                      len(hits), ops.ctypes.data, len(ops), lptr, lens, names, len(letters),
                                      int(secondary), buf.ctypes.data, cap, C.byref(out_len), C.byref(n_out))
        if st == 9 and out_len.value > cap:  # CAPACITY: grow to what it needs
            cap = out_len.value
            continue
        _check(st)
        return buf[:out_len.value].tobytes(), int(n_out.value)


def split_stream(stream: bytes) -> List[bytes]:
    """Records of a stream, each with its block_size."""
    out, p = [], 0
    while p < len(stream):
        bs = struct.unpack_from("<I", stream, p)[0]
        out.append(stream[p:p + 4 + bs])
        p += 4 + bs
    return out


def coordinate_sort(stream: bytes, n_refs: int) -> bytes:
    """--sort coordinate (align.rs:2418-2600): records bucketed by
    reference in header order, each bucket stably ordered by (position,
    reverse), then the unmapped records in input order."""
    recs = split_stream(stream)
    keyed = []
    for k, r in enumerate(recs):
        tid, pos = struct.unpack_from("<ii", r, 4)
        rev = struct.unpack_from("<H", r, 18)[0] & FLAG_REVERSE
        keyed.append(((tid, pos, 1 if rev else 0, k) if tid >= 0 else (n_refs, 0, 0, k), r))
    keyed.sort(key=lambda x: x[0])
    return b"".join(r for _, r in keyed)


def fnv1a64(data: bytes) -> int:
    """FNV-1a, 64-bit (align_unpaired.rs:56-60), over bytes."""
    h = 0xCBF29CE484222325
    for b in data:
        h = ((h ^ b) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


@dataclass
class RunOptions:
    """The unpaired CLI options that change the records (align.rs:134-289)."""
    mode: str = "local"
    scoring: Tuple[int, int, int, int] = (2, -1, -10, -1)
    min_score: int = 0
    max_ties: Optional[int] = None
    both_strands: bool = False
    max_read_len: int = 1000
    secondary: bool = False
    sort_coordinate: bool = False


Mapper = Callable[[List[np.ndarray], RunOptions], Tuple[np.ndarray, np.ndarray, np.ndarray]]


def run_unpaired(bam: BamFile, panel: Panel, opts: RunOptions, mapper: Mapper) -> Tuple[bytes, int]:
    """The output record stream of `eschalign READS -r REF -o OUT` for one
    input BAM: primary records only, reads over max_read_len written
    unmapped and never scored (their query is empty), the rest mapped by
    `mapper(codes, opts)` -> (maps, hits, ops), then build_records, then
    the coordinate sort when asked. Returns (stream, records)."""
    recs = primary_records(bam.records)
    codes = []
    for r in recs:
        letters = record_letters(r)
        over = opts.max_read_len != 0 and len(letters) > opts.max_read_len
        codes.append(np.zeros(0, np.uint8) if over else encode(letters))
    maps, hits, ops = mapper(codes, opts)
    stream, n = build_records(recs, maps, hits, ops, panel, opts.secondary)
    if opts.sort_coordinate:
        stream = coordinate_sort(stream, len(panel))
    return stream, n


def device_mapper(scorers: dict) -> Mapper:
    """A mapper over esch_map_reads: one GpuScorer per (mode, scoring),
    built on first use from the panel in scorers["panel"]."""
    from . import GpuScorer, Mode, Scoring

    def run(codes, opts):
        key = (opts.mode, opts.scoring)
        if key not in scorers:
            scorers[key] = GpuScorer(scorers["panel"], Scoring(*opts.scoring), Mode.parse(opts.mode))
        recs, hits, ops, _ = scorers[key].map_reads_raw(codes, both_strands=opts.both_strands,
                                                        min_score=opts.min_score, max_ties=opts.max_ties)
        return recs, hits, ops

    return run
