"""The unpaired run end to end on an MI355X: the fixture BAM's reads scored,
selected and traced by esch_map_reads on the device, the records built by
esch_bam_build_records, and the record stream's count and FNV-1a 64 digest
equal to the reference's own golden digests for every flag set of
crates/eschalign-cli/tests/align_unpaired.rs:21-36 (the CPU path's
counterpart is tests/test_bam_records.py)."""
import os

import pytest

import eschalign_amd as ea
from eschalign_amd import bam as B
from test_bam_records import CONFIGS, GOLDEN, golden

pytestmark = pytest.mark.gpu


def test_device_record_streams_equal_the_reference_digests():
    p = ea.Panel(B.read_fasta(os.path.join(GOLDEN, "trna_reference.fa")))
    bam = B.read_bam(os.path.join(GOLDEN, "trna_mappings_padded.bam"))
    mapper = B.device_mapper({"panel": p})
    want = golden()
    for label, opts in CONFIGS:
        stream, n = B.run_unpaired(bam, p, opts, mapper)
        assert (n, f"{B.fnv1a64(stream):016x}") == want[label], label
