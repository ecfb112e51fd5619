"""The host read-storage construction (rabbitqcplus_amd/reads.py) against the reference's
FASTQ -> storage construction (constructChunkedReadStorageFromFiles, compiled from source),
on reads with N and other non-ACGT characters, lower-case bases and more than one parser
batch of 65,536 reads (CPU only)."""
import numpy as np
import pytest

import oracle
from rabbitqcplus_amd import reads
from tests import fastq_data

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


@pytest.mark.parametrize("paired", [False, True])
def test_host_storage_matches_reference_parser(tmp_path, paired):
    data = fastq_data.make(seed=5 + paired, n_reads=70_000 if not paired else 2 * 36_000, paired=paired)
    files = fastq_data.write(tmp_path, data, paired)
    ref = oracle.ref_parse_fastq(*files, use_qualities=True)
    rs = reads.make_storage(data["seq_matrix"], data["lengths"], data["qual_matrix"], paired)
    assert ref["enc"].shape[0] == rs.n
    assert np.array_equal(ref["lengths"], rs.lengths)
    assert np.array_equal(ref["ambiguous"], rs.ambiguous)
    assert ref["ambiguous"].sum() > 100
    w = min(ref["enc"].shape[1], rs.enc.shape[1])
    assert np.array_equal(ref["enc"][:, :w], rs.enc[:, :w])
    # whole rows: the reference's storage holds zeros past each read's end, and the
    # correction reads them (reverse-complement candidates reverse the whole row)
    ml = ref["max_length"]
    assert np.array_equal(ref["quals"], rs.quals[:, :ml])
