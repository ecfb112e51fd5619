"""Candidate correction (--candidateCorrection) in the restatement (oracle/carecpu.cpp)
against the reference compiled from source (CPU only).

The reference keeps a candidate's record unless the candidate was corrected as an HQ
anchor before (include/corrector.hpp:1676-1680), which with several threads depends on
their timing; with one thread reads are processed in ascending order. The restatement
fixes that order, so it is compared with one-thread runs: the records in the order the
reference inserts them, and (golden, tests/golden/candidates.npz) sorted by read id.
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import oracle
from oracle import output_merge
from rabbitqcplus_amd import reads, sim
from tests import helpers
from tests.golden import make_golden as mg
from tests.test_output_oracle import sim_texts

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "candidates.npz"))


def cpu_records(rs, o, new_columns=15):
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_candidate_correction(True, new_columns)
    n, data = cpu.correct(threads=8)
    cpu.close()
    return n, data


@pytest.mark.parametrize("name", mg.CAND_CASES)
def test_candidate_records_match_golden(name):
    s, rs, o = helpers.case(name)
    assert mg.storage_digest(rs) == str(GOLDEN[f"{name}_digest"])
    n, data = cpu_records(rs, o)
    assert n == int(GOLDEN[f"{name}_num_records"])
    srt = oracle.sort_records(data)
    assert hashlib.sha256(srt).hexdigest() == str(GOLDEN[f"{name}_sha256"])
    if f"{name}_records" in GOLDEN:
        assert srt == GOLDEN[f"{name}_records"].tobytes()
    cands = sum(1 for r in output_merge.decode_records(data) if r["type"] == 1)
    assert cands == int(GOLDEN[f"{name}_num_candidate_records"]) > 0


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name, seed, new_columns", [("se_higherr", 401, 15), ("pe", 402, 15), ("varlen_q", 403, 5)])
def test_insertion_order_matches_reference_live(name, seed, new_columns):
    """Fresh seeds, the reference's own insertion order (one thread), and a narrower
    candidate window."""
    simkw, opts = helpers.CASES[name]
    s = sim.simulate(**dict(simkw, seed=seed))
    use_q = bool(opts.get("use_quality_scores", 0))
    rs = reads.storage_from_sim(s, use_qualities=use_q)
    o = dict(opts, paired=int(rs.paired), use_quality_scores=int(use_q))
    ref = oracle.CareRef(rs, threads=8, **o)
    ref.set_candidate_correction(True, new_columns)
    ref.set_record_order(False)
    nr, dr = ref.correct(threads=1)
    ref.close()
    nc, dc = cpu_records(rs, o, new_columns)
    assert nr == nc and dr == dc


def test_candidate_correction_off_is_unchanged():
    s, rs, o = helpers.case("pe")
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", "records_pe.npz"))["records"].tobytes()
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_candidate_correction(False, 15)
    assert cpu.correct(threads=8)[1] == golden
    cpu.close()


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("labels", [False, True])
def test_merge_of_candidate_records_matches_reference(labels):
    """Reads with an anchor record and several candidate records reach the combine rule's
    agreement branch; the merge restatement against the reference's merge."""
    texts, paired = sim_texts("varlen_q")
    records = GOLDEN["varlen_q_records"].tobytes()
    recs = output_merge.decode_records(records)
    per_read = np.bincount([r["read_id"] for r in recs])
    assert (per_read >= 3).sum() > 100
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "in.fq")
        open(path, "wb").write(texts[0])
        ref, changed = oracle.ref_output_fastq(path, None, records, labels=labels, tmpdir=d)
    assert changed == len(recs)
    ours = output_merge.merge(texts, records, paired, labels)
    assert ref == ours
    if labels:
        heads = ours[0].split(b"\n")[0::4]
        assert any(h.endswith(b"care:q=2") for h in heads)
