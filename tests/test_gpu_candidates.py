"""Candidate correction on the GPU (care_opts.correct_candidates).

Bar: bit-exact. The device records, in the order a one-thread run of the reference inserts
them, equal the restatement's (oracle/carecpu.cpp, pinned to the reference in
tests/test_candidates_oracle.py); sorted by read id they equal the reference's golden
digests; batch sizes and ranges that split 16-read groups do not change them; the merged
FASTQ of a run with candidate records equals the pinned merge restatement.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from oracle import output_merge
from rabbitqcplus_amd import engine
from tests import helpers
from tests.golden import make_golden as mg
from tests.test_output_oracle import sim_texts

pytestmark = pytest.mark.gpu

GOLDEN = np.load(os.path.join(os.path.dirname(__file__), "golden", "candidates.npz"))


def gpu_records(rs, o, first=0, n=None, **extra):
    eng = engine.CareEngine(0, correct_candidates=1, **dict(o, **extra))
    try:
        eng.set_reads(rs)
        eng.build_tables()
        return eng.correct(first, n)
    finally:
        eng.close()


def cpu_records(rs, o, first=0, n=None):
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_candidate_correction(True, 15)
    out = cpu.correct(first, n, threads=8)
    cpu.close()
    return out


@pytest.mark.parametrize("name", mg.CAND_CASES)
def test_candidate_records_bit_exact(name):
    s, rs, o = helpers.case(name)
    print(f"[candidates {name}] reads={rs.n}", flush=True)
    n, data = gpu_records(rs, o)
    nc, dc = cpu_records(rs, o)
    assert n == nc and data == dc
    assert n == int(GOLDEN[f"{name}_num_records"])
    assert hashlib.sha256(oracle.sort_records(data)).hexdigest() == str(GOLDEN[f"{name}_sha256"])


@pytest.mark.parametrize("name, batch, first, count", [
    ("se_higherr", 1000, 0, None),      # batches of 992 anchors
    ("se_higherr", 0, 5, 7001),         # a range that starts and ends inside 16-read groups
    ("pe", 2050, 6, 9000),              # paired, batches of 2048
])
def test_candidate_batches_and_ranges(name, batch, first, count):
    s, rs, o = helpers.case(name)
    count = rs.n - first if count is None else count
    print(f"[candidates {name} batch={batch} first={first} n={count}]", flush=True)
    n, data = gpu_records(rs, o, first, count, batch_size=batch)
    nc, dc = cpu_records(rs, o, first, count)
    assert n == nc and data == dc


def test_candidate_correction_fastq_to_fastq():
    s, rs, o = helpers.case("varlen_q")
    texts, paired = sim_texts("varlen_q")
    eng = engine.CareEngine(0, correct_candidates=1, **o)
    try:
        eng.set_reads_fastq(*texts)
        eng.build_tables()
        n, records = eng.correct()
        outs = [eng.write_fastq(records, 0, labels) for labels in (False, True)]
    finally:
        eng.close()
    assert oracle.sort_records(records) == GOLDEN["varlen_q_records"].tobytes()
    assert outs[0] == output_merge.merge(texts, records, False, False)[0]
    assert outs[1] == output_merge.merge(texts, records, False, True)[0]


def test_candidate_correction_limits():
    s, rs, o = helpers.case("se")
    eng = engine.CareEngine(0, correct_candidates=1, engine=2, **o)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        with pytest.raises(engine.CareError, match="staged engine"):
            eng.correct()
    finally:
        eng.close()
    s, rs, o = helpers.case("len400")
    eng = engine.CareEngine(0, correct_candidates=1, **o)
    try:
        eng.set_reads(rs)
        eng.build_tables()
        with pytest.raises(engine.CareError, match="340"):
            eng.correct()
    finally:
        eng.close()
