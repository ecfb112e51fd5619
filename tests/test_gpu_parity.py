"""GPU parity: every stage of the HIP path against the CPU oracle on the same seeded reads.

Bar: bit-exact signatures, candidate ids, alignment (score, overlap, shift, nOps,
validity, orientation), MSA counts/coverage/consensus and records; MSA float weights
and supports within 1e-5 relative (observed: identical).
"""
import numpy as np
import pytest

import oracle
from rabbitqcplus_amd import sim, engine
from tests import helpers

pytestmark = pytest.mark.gpu

WEIGHT_RTOL = 1e-5


def _engine(name, **extra):
    s, rs, o = helpers.case(name)
    o = dict(o)
    o.update(extra)
    eng = engine.CareEngine(0, **o)
    eng.set_reads(rs)
    eng.build_tables()
    return s, rs, o, eng


@pytest.fixture(scope="module")
def se_gpu():
    s, rs, o, eng = _engine("se", keep_intermediates=1, batch_size=4096)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    yield s, rs, o, eng, cpu
    eng.close()
    cpu.close()


def test_signatures_bit_exact(se_gpu):
    s, rs, o, eng, cpu = se_gpu
    g = eng.signatures()
    c = cpu.signatures()
    assert g.shape == c.shape
    assert np.array_equal(g, c)


def test_candidates_alignments_msa_first_batch(se_gpu):
    s, rs, o, eng, cpu = se_gpu
    n = 4096
    eng.correct(0, n)
    ins = eng.inspect_batch()
    assert ins.first == 0 and ins.n == n
    rng = np.random.default_rng(0)
    sample = set(rng.choice(n, 300, replace=False).tolist()) | {0, n - 1}
    for i in range(n):
        g = ins.candidate_ids[ins.offsets[i]:ins.offsets[i + 1]]
        c = cpu.candidates(i)
        assert np.array_equal(g, c), f"candidates of anchor {i}"
    for i in sorted(sample):
        tr = cpu.trace(i)
        b, e = ins.offsets[i], ins.offsets[i + 1]
        al = ins.alignments[b:e]
        assert np.array_equal(al[:, 0:5], tr["fwd_alignments"]), f"fwd alignments of anchor {i}"
        assert np.array_equal(al[:, 5:10], tr["rc_alignments"]), f"rc alignments of anchor {i}"
        assert np.array_equal(al[:, 10], tr["orientation"].astype(np.int32)), f"orientation of anchor {i}"
        info = ins.anchor_info[i]
        assert info[0] == tr["num_filtered"], f"filtered count of anchor {i}"
        assert info[1] == tr["num_final"], f"final count of anchor {i}"
        if tr["num_final"] == 0:
            continue
        assert info[2] == tr["num_columns"] and info[3] == tr["anchor_column_begin"]
        assert info[5] == tr["is_corrected"] and info[4] == tr["is_hq"]
        assert info[6] == tr["min_coverage"] and info[7] == tr["max_coverage"]
        assert np.float32(info[9:10].view(np.float32)[0]) == np.float32(tr["avg_support"])
        assert np.float32(info[10:11].view(np.float32)[0]) == np.float32(tr["min_support"])
        m = eng.inspect_msa(i)
        assert m["num_columns"] == tr["num_columns"]
        assert np.array_equal(m["counts"], tr["counts"])
        assert np.array_equal(m["coverage"], tr["coverage"])
        assert np.array_equal(m["consensus"], tr["consensus"])
        np.testing.assert_allclose(m["weights"], tr["weights"], rtol=WEIGHT_RTOL, atol=0)
        np.testing.assert_allclose(m["support"], tr["support"], rtol=WEIGHT_RTOL, atol=0)


@pytest.mark.parametrize("name", list(helpers.CASES))
def test_records_match_oracle(name):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n} opts={o}", flush=True)
    n_gpu, d_gpu = engine.correct_gpu(rs, **o)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    n_cpu, d_cpu = cpu.correct(threads=8)
    cpu.close()
    assert n_gpu == n_cpu
    if d_gpu != d_cpu:
        rg = {r[0]: r for r in helpers.parse_records(d_gpu)}
        rc = {r[0]: r for r in helpers.parse_records(d_cpu)}
        diff = sorted(k for k in set(rg) | set(rc) if rg.get(k) != rc.get(k))
        pytest.fail(f"{len(diff)} differing records, first ids {diff[:10]}")
    if oracle.ref_available():
        ref = oracle.CareRef(rs, threads=8, **o)
        n_ref, d_ref = ref.correct(threads=8)
        ref.close()
        assert d_gpu == d_ref


STAGED_CASES = ["se_q", "pe", "varlen_q", "nbases_exclude", "se_tandem", "len300_q"]


@pytest.mark.parametrize("name", STAGED_CASES)
def test_staged_engine_matches_fused_and_oracle(name):
    """The staged per-stage kernels (engine=1) and the persistent fused kernel (engine=2)
    give the same record bytes, equal to the oracle's."""
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    _, d_staged = engine.correct_gpu(rs, engine=1, **o)
    _, d_fused = engine.correct_gpu(rs, engine=2, **o)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    _, d_cpu = cpu.correct(threads=8)
    cpu.close()
    assert d_staged == d_cpu
    assert d_fused == d_cpu


def test_fused_engine_rejects_wide_reads():
    s, rs, o = helpers.case("len400")
    eng = engine.CareEngine(0, engine=2, **o)
    eng.set_reads(rs)
    eng.build_tables()
    with pytest.raises(engine.CareError, match="too long for the fused engine"):
        eng.correct()
    eng.close()


@pytest.mark.parametrize("paired", [False, True])
def test_fused_range_split_invariance(paired):
    """Fused records of [0, n) equal those of [0, m) followed by [m, n), also with
    fewer resident waves than work units."""
    s, rs, o = helpers.case("pe" if paired else "se_higherr")
    eng = engine.CareEngine(0, engine=2, **o)
    eng.set_reads(rs)
    eng.build_tables()
    whole = eng.correct()[1]
    m = 1234
    assert eng.correct(0, m)[1] + eng.correct(m, rs.n - m)[1] == whole
    eng.close()
    eng = engine.CareEngine(0, engine=2, waves_per_cu=1, **o)
    eng.set_reads(rs)
    eng.build_tables()
    assert eng.correct()[1] == whole
    eng.close()


def test_batching_invariance():
    """Staged record bytes do not depend on the batch size or on the anchor range split."""
    s, rs, o = helpers.case("se_higherr")
    outs = []
    for bs in (0, 1000, 4097):
        eng = engine.CareEngine(0, batch_size=bs, engine=1, **o)
        eng.set_reads(rs)
        eng.build_tables()
        outs.append(eng.correct()[1])
        if bs == 1000:
            a = eng.correct(0, 1234)[1]
            b = eng.correct(1234, rs.n - 1234)[1]
            assert a + b == outs[-1]
        eng.close()
    assert outs[0] == outs[1] == outs[2]


@pytest.mark.parametrize("name", ["se_q", "varlen_q", "se_higherr"])
def test_msa_tile_invariance(name):
    """The number of candidates the one-wavefront MSA stages into LDS per step (4..64) does
    not change the records; they equal the oracle's."""
    s, rs, o = helpers.case(name)
    print(f"[case {name}] msa tiles", flush=True)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    _, d_cpu = cpu.correct(threads=8)
    cpu.close()
    for tile in (4, 12, 64):
        _, d = engine.correct_gpu(rs, engine=1, msa_tile=tile, **o)
        assert d == d_cpu, tile


def test_correction_reduces_error_rate():
    s, rs, o = helpers.case("se_higherr")
    _, data = engine.correct_gpu(rs, **o)
    before = helpers.error_rate(s, [rs.decoded(i) for i in range(rs.n)])
    after = helpers.error_rate(s, helpers.apply_records(rs, data))
    print(f"error rate {before:.5f} -> {after:.5f}")
    assert after < 0.25 * before


@pytest.mark.parametrize("name", ["se_higherr_q", "pe"])
def test_fused_graph_replays_match_eager(name):
    """The fused pass captured into a HIP graph and replayed: every replay equals the
    eager pass, with eager work on another context between replays; a graph is reused
    while the pass's arguments hold and recaptured when the anchor range changes."""
    s, rs, o = helpers.case(name)
    eager = engine.CareEngine(0, engine=2, **o)
    eager.set_reads(rs)
    eager.build_tables()
    want = eager.correct()[1]
    g = engine.CareEngine(0, engine=2, use_graphs=1, **o)
    g.set_reads(rs)
    g.build_tables()
    for rep in range(3):
        assert g.correct()[1] == want, f"replay {rep}"
        assert eager.correct(0, 2000)[1] == eager.correct(0, 2000)[1]
    assert g.timings()["graphs_built"] == 1
    m = 1234
    assert g.correct(0, m)[1] + g.correct(m, rs.n - m)[1] == want
    assert g.timings()["graphs_built"] == 3
    assert g.correct()[1] == want
    g.close()
    eager.close()


def test_fused_graph_recaptured_after_the_output_merge_reallocates():
    """Same context: a graph-captured fused pass, then the output merge (its radix sort
    can grow the context's scan workspace, which the captured record scan uses), then the
    pass again. The second pass is recaptured when a buffer it uses moved, and its records
    equal the first pass's and the eager engine's."""
    s, rs, o = helpers.case("se_higherr_q")
    eager = engine.CareEngine(0, engine=1, **o)
    eager.set_reads(rs)
    eager.build_tables()
    want = eager.correct()[1]
    eager.close()
    g = engine.CareEngine(0, engine=2, use_graphs=1, **o)
    try:
        g.set_reads_fastq(*sim.fastq_texts(s))
        g.build_tables()
        first = g.correct()[1]
        assert first == want
        built = g.timings()["graphs_built"]
        out = g.write_fastq(first, 0)
        assert len(out) > 0
        for _ in range(2):
            assert g.correct()[1] == want
        assert g.timings()["graphs_built"] in (built, built + 1)
    finally:
        g.close()


def test_autotuner_choice_is_exact(tmp_path):
    """The autotuner times every schedule on a sample and caches its choice; the chosen
    schedule corrects the whole read set exactly like the oracle."""
    from rabbitqcplus_amd import autotune
    s, rs, o = helpers.case("se_higherr_q")
    cache = str(tmp_path / "tune.json")
    r = autotune.tune(rs, sample=4096, reps=2, cache=cache, **o)
    cands = autotune.candidates(o)
    assert not r["cached"] and r["choice"] in cands
    assert len(r["timings_ms"]) == len(cands)
    # every schedule, the tiles swept included, gives the oracle's records
    for c in cands[::3]:
        assert engine.correct_gpu(rs, **{**o, **c})[1] == engine.correct_gpu(rs, **o)[1], c
    assert autotune.tune(rs, sample=4096, reps=2, cache=cache, **o)["cached"]
    _, d = engine.correct_gpu(rs, **{**o, **r["choice"]})
    cpu = oracle.CareCpu(rs, threads=8, **o)
    _, d_cpu = cpu.correct(threads=8)
    cpu.close()
    assert d == d_cpu
