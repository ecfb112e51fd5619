"""GPU tests of the seeded alignment (care_opts.seeded_alignment), the fast learned form's
alignment: one shift per candidate, placed by the seed of the first minhash map the
candidate shares with its anchor (an addition; the reference has no counterpart).

Bar: records bit-exact against the oracle's seeded mode (tests/test_seeded_oracle.py pins
that mode by brute force), with the classic decision and with the learned seam; the
learned form on seeded alignment leaves no more errors than the plain (reference) path on
the same reads.
"""
import numpy as np
import pytest

import oracle
from rabbitqcplus_amd import engine, evaluate, learned
from tests import helpers

pytestmark = pytest.mark.gpu

CASES = ["se_q", "se_higherr_q", "se_repeats", "pe", "varlen_q", "nbases", "short"]


@pytest.mark.parametrize("name", CASES)
def test_seeded_records_bit_exact(name):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    n_g, d_g = engine.correct_gpu(rs, seeded_alignment=1, **o)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_seeded_alignment(True)
    n_c, d_c = cpu.correct(threads=8)
    cpu.close()
    assert n_g == n_c and d_g == d_c


@pytest.mark.parametrize("name", ["se_higherr_q", "se_repeats", "pe", "varlen_q"])
def test_seeded_learned_records_bit_exact(name):
    s, rs, o = helpers.case(name)
    params, _ = learned.load_classifier()
    print(f"[case {name}] reads={rs.n}", flush=True)
    n_g, d_g = engine.correct_gpu(rs, classifier=params, seeded_alignment=1, **o)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_seeded_alignment(True)
    cpu.set_classifier(params)
    n_c, d_c = cpu.correct(threads=8)
    cpu.close()
    assert n_g == n_c and d_g == d_c


def test_seeded_alignment_rejects_unsupported_configurations():
    s, rs, o = helpers.case("len400")
    eng = engine.CareEngine(0, seeded_alignment=1, **o)
    eng.set_reads(rs)
    eng.build_tables()
    with pytest.raises(engine.CareError, match="352"):
        eng.correct()
    eng.close()
    s, rs, o = helpers.case("se")
    eng = engine.CareEngine(0, seeded_alignment=1, engine=2, **o)
    eng.set_reads(rs)
    eng.build_tables()
    with pytest.raises(engine.CareError, match="staged"):
        eng.correct()
    eng.close()


@pytest.mark.parametrize("name", ["se_higherr_q", "se_repeats", "varlen_q"])
def test_seeded_learned_error_rate_not_above_plain(name):
    """Errors left after correction (against the simulator's truth): the learned form on
    seeded alignment leaves at most as many as the plain path."""
    s, rs, o = helpers.case(name)
    params, _ = learned.load_classifier()
    _, d_plain = engine.correct_gpu(rs, **o)
    _, d_fast = engine.correct_gpu(rs, classifier=params, seeded_alignment=1, **o)
    ep = evaluate.error_counts(s, rs, d_plain)
    ef = evaluate.error_counts(s, rs, d_fast)
    print(f"[{name}] plain {ep} / seeded learned {ef}", flush=True)
    assert ef["after"] <= ep["after"]
