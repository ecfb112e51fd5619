"""The learned seam's oracle, pinned (CPU only).

- decision columns and the 13 reference features: against the reference's feature print
  mode (golden sample + digest; live when oracle/_ref is built);
- the seam itself: with a classifier whose logit is always 0 the records equal the
  reference's classic records; with logit +100 / -100 they equal the reference's Forest
  mode with a forest that always answers 1 / 0;
- the double-precision gradient: central finite differences;
- the bfloat16 inference arithmetic: torch's bfloat16 rounding and an exactly
  representable network evaluated in float64.
"""
import hashlib
import os
import tempfile

import numpy as np
import pytest

import oracle
from tests import helpers
from tests.golden import make_golden as mg

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "seam.npz")
P = oracle.MLP_PARAMS
OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 1024, 1088, 5184, 5248, 5312


def _sorted_key(meta):
    return meta[:, 0].astype(np.int64), meta[:, 1].astype(np.int64), (meta[:, 2] >> 2) & 3


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name", mg.SEAM_CASES)
def test_decision_columns_and_features_golden(golden, name):
    s, rs, o = helpers.case(name)
    assert str(golden[f"{name}_digest"]) == mg.storage_digest(rs)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    f, m = cpu.features(threads=8)
    cpu.close()
    rid, pos, cn = _sorted_key(m)
    order = np.lexsort((pos, rid))
    key = np.stack([rid[order], pos[order], np.frombuffer(b"ACGT", np.uint8)[cn[order]].astype(np.int64)], 1)
    assert len(key) == int(golden[f"{name}_num_columns"])
    assert hashlib.sha256(key.astype("<i8").tobytes()).hexdigest() == str(golden[f"{name}_columns_sha256"])
    rows = golden[f"{name}_sample_rows"]
    assert np.array_equal(key[rows], golden[f"{name}_sample_key"])
    ours = f[order][rows, :13].astype(np.float64)
    ref = golden[f"{name}_sample_features"]
    # the reference prints six significant digits
    np.testing.assert_allclose(ours, ref, rtol=6e-6, atol=1e-12)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_features_live_reference_fresh_seed():
    s, rs, o = helpers.case("se_higherr")
    from rabbitqcplus_amd import reads, sim
    s = sim.simulate(genome_length=40_000, read_length_min=150, read_length_max=150, coverage=30.0,
                     q_first=26.0, q_last=14.0, repeat_fraction=0.1, repeat_unit_length=800, repeat_copies=6,
                     repeat_divergence=0.01, seed=4321)
    rs = reads.storage_from_sim(s, use_qualities=True)
    o = dict(use_quality_scores=1, paired=0)
    ref = oracle.CareRef(rs, threads=1, **o)
    with tempfile.TemporaryDirectory() as d:
        rid, pos, cons, feats = ref.print_anchor_features(os.path.join(d, "f.txt"))
    ref.close()
    cpu = oracle.CareCpu(rs, threads=8, **o)
    f, m = cpu.features(threads=8)
    cpu.close()
    order = np.lexsort((pos, rid))
    assert np.array_equal(rid[order], m[:, 0]) and np.array_equal(pos[order], m[:, 1])
    assert np.array_equal(np.array([ord(c) for c in cons[order]]), np.frombuffer(b"ACGT", np.uint8)[(m[:, 2] >> 2) & 3])
    np.testing.assert_allclose(f[:, :13], feats[order], rtol=6e-6, atol=1e-12)
    assert len(m) > 1000


def _learned_records(name, b3):
    s, rs, o = helpers.case(name)
    params = np.zeros(P, np.float32)
    params[OFF_B3] = b3
    cpu = oracle.CareCpu(rs, threads=8, **o)
    cpu.set_classifier(params)
    n, data = cpu.correct(threads=8)
    cpu.close()
    return n, data


@pytest.mark.parametrize("name", mg.SEAM_CASES)
def test_seam_extremes_golden(golden, name):
    # logit 0: every decision falls back to the classic rule -> the reference's records
    n0, d0 = _learned_records(name, 0.0)
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", f"records_{name}.npz"))
    assert d0 == g["records"].tobytes() and n0 == int(g["num_records"])
    for tag, b3 in (("yes", 100.0), ("no", -100.0)):
        n, d = _learned_records(name, b3)
        assert n == int(golden[f"{name}_forest_{tag}_records"]), tag
        assert hashlib.sha256(d).hexdigest() == str(golden[f"{name}_forest_{tag}_sha256"]), tag


def _random_params(rng, scale=0.5):
    p = rng.normal(0.0, scale, P)
    p[OFF_W1:OFF_B1] *= 1.0 / np.sqrt(16)
    p[OFF_W2:OFF_B2] *= 1.0 / np.sqrt(64)
    p[OFF_W3:OFF_B3] *= 1.0 / np.sqrt(64)
    return p


def test_mlp_gradient_finite_differences():
    rng = np.random.default_rng(3)
    params = _random_params(rng)
    X = rng.normal(0.5, 0.4, (48, 16)).astype(np.float32)
    y = (rng.random(48) < 0.6).astype(np.float32)
    loss, G, _ = oracle.mlp_loss_grad_f64(params, X, y)
    idx = np.concatenate([rng.choice(P, 300, replace=False), np.arange(OFF_W3, P)])
    h = 1e-6
    num = np.zeros(len(idx))
    for t, i in enumerate(idx):
        pp = params.copy(); pp[i] += h
        pm = params.copy(); pm[i] -= h
        num[t] = (oracle.mlp_loss_grad_f64(pp, X, y, grad=False)[0] - oracle.mlp_loss_grad_f64(pm, X, y, grad=False)[0]) / (2 * h)
    err = np.abs(num - G[idx])
    assert err.max() <= 1e-7 * max(1.0, np.abs(G).max()), err.max()
    # a planted mistake in the analytic gradient would show: the check is not vacuous
    assert np.abs(G[idx]).max() > 1e-3


def test_mlp_loss_closed_form():
    """Zero weights: every logit is b3, loss = mean softplus(b3) - y*b3, dL/db3 = sigmoid(b3) - mean(y)."""
    rng = np.random.default_rng(4)
    X = rng.normal(0, 1, (100, 16)).astype(np.float32)
    y = (rng.random(100) < 0.3).astype(np.float32)
    params = np.zeros(P)
    params[OFF_B3] = 0.7
    loss, G, z = oracle.mlp_loss_grad_f64(params, X, y)
    assert np.all(z == 0.7)
    y64 = y.astype(np.float64)
    assert loss == pytest.approx(np.mean(np.log1p(np.exp(0.7)) - y64 * 0.7), rel=1e-12)
    assert G[OFF_B3] == pytest.approx(1 / (1 + np.exp(-0.7)) - y64.mean(), rel=1e-12)
    assert np.all(G[:OFF_B3] == 0)


def test_bf16_rounding_matches_torch():
    """A network that passes feature 0 through (W1[0,0] = W2[0,0] = W3[0] = 1) returns
    bf16(x0) for x0 > 0: the rounding is torch's round-to-nearest-even bfloat16."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    params = np.zeros(P, np.float32)
    params[OFF_W1 + 0] = 1.0
    params[OFF_W2 + 0] = 1.0
    params[OFF_W3 + 0] = 1.0
    x0 = np.concatenate([rng.uniform(0.001, 4.0, 5000),
                         # exact halfway cases between bfloat16 neighbours
                         (np.arange(1, 200) * 2 + 1).astype(np.float64) * 2.0 ** -9]).astype(np.float32)
    X = np.zeros((len(x0), 16), np.float32)
    X[:, 0] = x0
    z = oracle.mlp_logits_bf16(params, X)
    want = torch.from_numpy(x0).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(z, want)
    assert not np.array_equal(z, x0)


def test_exact_network_matches_float64():
    """Sparse weights in {-1, -0.5, 0.5, 1} and inputs in multiples of 1/8 keep every
    sum exact in bfloat16/float, so the device arithmetic equals a float64 forward pass
    with the parameter layout W1[64][16], b1, W2[64][64], b2, W3, b3."""
    rng = np.random.default_rng(6)
    vals = np.array([-1.0, -0.5, 0.5, 1.0])
    params = np.zeros(P, np.float64)
    W1 = np.zeros((64, 16)); W2 = np.zeros((64, 64)); W3 = np.zeros(64)
    for j in range(64):
        W1[j, rng.choice(16, 2, replace=False)] = rng.choice(vals, 2)
        W2[j, rng.choice(64, 2, replace=False)] = rng.choice(vals, 2)
    W3[rng.choice(64, 4, replace=False)] = rng.choice(vals, 4)
    b1 = rng.choice([-0.25, 0.0, 0.25], 64); b2 = rng.choice([-0.25, 0.0, 0.25], 64); b3 = 0.125
    params[OFF_W1:OFF_B1] = W1.ravel(); params[OFF_B1:OFF_W2] = b1
    params[OFF_W2:OFF_B2] = W2.ravel(); params[OFF_B2:OFF_W3] = b2
    params[OFF_W3:OFF_B3] = W3; params[OFF_B3] = b3
    X = (rng.integers(-4, 5, (400, 16)) / 8.0).astype(np.float32)
    h1 = np.maximum(X.astype(np.float64) @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    want = h2 @ W3 + b3
    z = oracle.mlp_logits_bf16(params.astype(np.float32), X)
    assert np.array_equal(z.astype(np.float64), want)
    assert np.unique(want).size > 10
    # and the double-precision path agrees
    _, _, z64 = oracle.mlp_loss_grad_f64(params, X, None, grad=False)
    assert np.array_equal(z64, want)
