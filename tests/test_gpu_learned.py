"""GPU tests of the learned seam and its training.

Bar: decision columns, their 16 features and the learned records are bit-exact against
the oracle (which emulates the device's bfloat16 inference arithmetic); the MLP gradient
of the device kernels agrees with the oracle's double-precision backpropagation (itself
pinned by finite differences) within 1e-4 of the largest component; one Adam step equals
its closed form; a classifier trained on the GPU lowers the post-correction error rate of
held-out reads below the classic decision's.
"""
import hashlib
import os

import numpy as np
import pytest

import oracle
from rabbitqcplus_amd import engine, evaluate, learned, reads, sim
from tests import helpers
from tests.golden import make_golden as mg

pytestmark = pytest.mark.gpu

FEATURE_CASES = ["se_higherr", "se_higherr_q", "se_repeats", "pe", "varlen_q", "nbases"]
OFF_B3 = learned.OFF_B3


def _random_classifier(seed):
    """A random classifier whose logits spread over both confidence thresholds on the
    test cases, so that all three outcomes (take, keep, fall back) occur."""
    p = learned.init_params(seed)
    p[learned.OFF_W3:OFF_B3] *= 20.0
    p[OFF_B3] = 52.5
    return p


@pytest.mark.parametrize("name", FEATURE_CASES)
def test_features_bit_exact(name):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    eng = engine.CareEngine(0, **o)
    eng.set_reads(rs)
    eng.build_tables()
    fg, mg_ = eng.extract_features()
    eng.close()
    cpu = oracle.CareCpu(rs, threads=8, **o)
    fc, mc = cpu.features(threads=8)
    cpu.close()
    assert len(mc) > 0
    assert np.array_equal(mg_, mc)
    assert np.array_equal(fg.view(np.uint32), fc.view(np.uint32))


@pytest.mark.parametrize("name", ["se_higherr_q", "se_repeats", "pe", "varlen_q"])
@pytest.mark.parametrize("eng_kind", [1, 2])
def test_learned_records_bit_exact(name, eng_kind):
    s, rs, o = helpers.case(name)
    params = _random_classifier(7)
    print(f"[case {name} engine {eng_kind}] reads={rs.n}", flush=True)
    n_g, d_g = engine.correct_gpu(rs, classifier=params, engine=eng_kind, **o)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_classifier(params)
    n_c, d_c = cpu.correct(threads=8)
    f, m = cpu.features(threads=8)
    cpu.close()
    assert n_g == n_c and d_g == d_c
    # the classifier takes, keeps and leaves undecided some decision columns
    z = oracle.mlp_logits_bf16(params, f)
    assert (z >= 2.1972245773).any() and (z <= -2.1972245773).any() and (np.abs(z) < 2.1972245773).any()


@pytest.mark.parametrize("name", mg.SEAM_CASES)
def test_seam_extremes_gpu(name):
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", "seam.npz"))
    s, rs, o = helpers.case(name)
    rec = np.load(os.path.join(os.path.dirname(__file__), "golden", f"records_{name}.npz"))
    for tag, b3 in (("classic", 0.0), ("yes", 100.0), ("no", -100.0)):
        p = np.zeros(learned.N_PARAMS, np.float32)
        p[OFF_B3] = b3
        n, d = engine.correct_gpu(rs, classifier=p, **o)
        if tag == "classic":
            assert d == rec["records"].tobytes()
        else:
            assert n == int(golden[f"{name}_forest_{tag}_records"])
            assert hashlib.sha256(d).hexdigest() == str(golden[f"{name}_forest_{tag}_sha256"])


def test_learned_needs_classifier_and_short_reads():
    s, rs, o = helpers.case("se")
    eng = engine.CareEngine(0, correction_type=1, **o)
    eng.set_reads(rs)
    eng.build_tables()
    with pytest.raises(engine.CareError, match="set_classifier"):
        eng.correct()
    with pytest.raises(engine.CareError, match="parameters"):
        eng.set_classifier(np.zeros(10, np.float32))
    eng.close()
    s, rs, o = helpers.case("len400")
    eng = engine.CareEngine(0, correction_type=1, **o)
    eng.set_reads(rs)
    eng.build_tables()
    eng.set_classifier(np.zeros(learned.N_PARAMS, np.float32))
    with pytest.raises(engine.CareError, match="340"):
        eng.correct()
    eng.close()


def _grad_case(seed, n):
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    p[learned.OFF_B1:learned.OFF_W2] = rng.normal(0, 0.1, 64)
    p[learned.OFF_B2:learned.OFF_W3] = rng.normal(0, 0.1, 64)
    X = rng.normal(0, 1, (n, 16)).astype(np.float32)
    y = (rng.random(n) < 0.5).astype(np.float32)
    return p, X, y


@pytest.mark.parametrize("n", [1, 63, 1000, 65536])
def test_forward_backward_vs_cpu_double(n):
    p, X, y = _grad_case(n, n)
    mlp = learned.Mlp(0)
    mlp.set_params(p)
    z = mlp.forward(X)
    loss, g = mlp.loss_grad(X, y)
    mlp.close()
    sub = slice(0, min(n, 4096))
    loss64, g64, z64 = oracle.mlp_loss_grad_f64(p.astype(np.float64), X, y)
    np.testing.assert_allclose(z[sub], z64[sub], rtol=1e-4, atol=1e-5)
    assert abs(loss - loss64) <= 1e-5 * max(1.0, abs(loss64))
    scale = np.abs(g64).max()
    assert scale > 0
    assert np.abs(g - g64).max() <= 1e-4 * scale, np.abs(g - g64).max() / scale


def test_adam_first_step_closed_form():
    """With zero moments, Adam's first step is p - lr * g / (|g| + eps)."""
    p, X, y = _grad_case(3, 5000)
    mlp = learned.Mlp(0)
    mlp.set_params(p)
    _, g = mlp.loss_grad(X, y)
    mlp.set_data(X, y)
    lr = 1e-3
    mlp.train(epochs=1, batch=len(X), lr=lr, seed=0)
    p1 = mlp.get_params()
    mlp.close()
    want = p.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1, want, rtol=0, atol=2e-6)


def test_training_reduces_loss_and_is_deterministic():
    rng = np.random.default_rng(11)
    X = rng.normal(0, 1, (20000, 16)).astype(np.float32)
    w = rng.normal(0, 1, 16)
    y = ((X @ w + 0.5 * np.sin(3 * X[:, 0])) > 0).astype(np.float32)
    runs = []
    for _ in range(2):
        mlp = learned.Mlp(0)
        mlp.set_params(learned.init_params(1))
        mlp.set_data(X, y)
        losses = mlp.train(epochs=20, batch=4096, lr=3e-3, seed=5)
        runs.append((losses, mlp.get_params()))
        mlp.close()
    losses = runs[0][0]
    assert losses[-1] < 0.35 * losses[0], losses
    assert np.array_equal(runs[0][1], runs[1][1])
    acc = ((oracle.mlp_loss_grad_f64(runs[0][1].astype(np.float64), X, y, grad=False)[2] > 0) == (y > 0.5)).mean()
    assert acc > 0.95


def _sets(seeds, profiles):
    out = []
    for name, kw, use_q in profiles:
        for seed in seeds:
            s = sim.simulate(seed=seed, **kw)
            rs = reads.storage_from_sim(s, use_qualities=use_q)
            out.append((name, s, rs, dict(use_quality_scores=int(use_q), paired=int(rs.paired))))
    return out


GAIN_PROFILES = [
    ("hierr", dict(genome_length=60_000, read_length_min=150, read_length_max=150, coverage=30.0,
                   q_first=24.0, q_last=16.0), True),
    ("rep10_hierr", dict(genome_length=80_000, repeat_fraction=0.15, repeat_unit_length=1000, repeat_copies=10,
                         repeat_divergence=0.01, read_length_min=150, read_length_max=150, coverage=30.0,
                         q_first=26.0, q_last=16.0), True),
]


def test_trained_classifier_gains_over_classic():
    """Train on the GPU from the device's own decision columns, then correct held-out
    reads: fewer post-correction errors than the classic decision; new miscorrections
    (bases right before, wrong after) stay a small fraction of the errors removed."""
    train = _sets((201, 202), GAIN_PROFILES)
    Xs, ys = [], []
    for name, s, rs, o in train:
        X, y, _ = learned.collect(s, rs, **o)
        Xs.append(X); ys.append(y)
    X = np.concatenate(Xs); y = np.concatenate(ys)
    params, losses, info = learned.train_classifier(X, y, epochs=40, batch=8192, lr=2e-3, seed=0)
    print(f"train rows {len(y)} loss {losses[0]:.4f} -> {losses[-1]:.4f}", flush=True)
    assert losses[-1] < losses[0]
    tot = dict(before=0, plain=0, learned=0, mis_plain=0, mis_learned=0)
    for name, s, rs, o in _sets((977,), GAIN_PROFILES):
        _, plain = engine.correct_gpu(rs, **o)
        _, learn = engine.correct_gpu(rs, classifier=params, **o)
        cp, cl = evaluate.error_counts(s, rs, plain), evaluate.error_counts(s, rs, learn)
        print(f"{name}: before {cp['before']} plain {cp['after']} learned {cl['after']} "
              f"miscorrected {cp['miscorrected']} / {cl['miscorrected']}", flush=True)
        tot["before"] += cp["before"]; tot["plain"] += cp["after"]; tot["learned"] += cl["after"]
        tot["mis_plain"] += cp["miscorrected"]; tot["mis_learned"] += cl["miscorrected"]
    assert tot["learned"] < tot["plain"], tot
    assert tot["mis_learned"] - tot["mis_plain"] <= 0.05 * (tot["plain"] - tot["learned"]), tot


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_CLASSIFIER), reason="no shipped classifier file")
def test_shipped_classifier_not_worse_on_benchmark_profile():
    params, info = learned.load_classifier()
    s = sim.workload("se_200kbp_150bp_30x", seed=31337)
    rs = reads.storage_from_sim(s, use_qualities=True)
    o = dict(use_quality_scores=1, paired=0)
    _, plain = engine.correct_gpu(rs, **o)
    _, learn = engine.correct_gpu(rs, classifier=params, **o)
    cp, cl = evaluate.error_counts(s, rs, plain), evaluate.error_counts(s, rs, learn)
    print(f"before {cp['before']} plain {cp['after']} learned {cl['after']}", flush=True)
    assert cl["after"] <= cp["after"]
