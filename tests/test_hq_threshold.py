"""The HQ test of the decision (src/msa.cpp:357-372) as the HIP kernels evaluate it: instead
of the loop over percent = pct..0 that records the smallest percent whose thresholds
1 - f and 1 - 3f (f = percent / 100) both supports meet and then asks whether it is at most
pct / 2, the kernels test the thresholds of hqPercent, the largest percent with
fleq(percent, pct / 2). Both forms are evaluated here in float32 over many error rates and
supports placed on, just above and just below every threshold the loop can use, and within
the 1e-5 tolerance band of fgeq (CPU only)."""
import numpy as np

f32 = np.float32


def feq(l, r):
    d = f32(l - r)
    return abs(d) < f32(1e-5)


def fgeq(l, r):
    return l > r or feq(l, r)


def fleq(l, r):
    return l < r or feq(l, r)


def thresholds(percent):
    factor = f32(f32(percent) / f32(100.0))
    return f32(f32(1.0) - f32(1.0) * factor), f32(f32(1.0) - f32(3.0) * factor)


def hq_loop(avg, mn, pct):
    smallest = 100
    for percent in range(pct, -1, -1):
        a, m = thresholds(percent)
        if fgeq(avg, a) and fgeq(mn, m):
            smallest = percent
    return fleq(f32(smallest), f32(pct) * f32(0.5))


def hq_single(avg, mn, pct):
    half = f32(pct) * f32(0.5)
    hq_percent = max(q for q in range(pct + 1) if fleq(f32(q), half))
    a, m = thresholds(hq_percent)
    return fgeq(avg, a) and fgeq(mn, m)


def probe_values(pct):
    vals = set()
    for percent in range(pct + 1):
        for t in thresholds(percent):
            for d in (0.0, 1e-5, 2e-5, 5e-6, 1e-6):
                for s in (1, -1):
                    v = f32(t + f32(s * d))
                    vals.update([v, np.nextafter(v, f32(2)), np.nextafter(v, f32(-2))])
    return sorted(vals)


def test_single_threshold_equals_percent_loop():
    rng = np.random.default_rng(5)
    checked = 0
    for e in [0.01, 0.02, 0.03, 0.05, 0.06, 0.07, 0.08, 0.1, 0.125, 0.15, 0.2]:
        pct = int(np.ceil(f32(e) * f32(100.0)))
        vals = probe_values(pct)
        extra = [f32(x) for x in rng.uniform(0.3, 1.0, 300)]
        pool = vals + extra
        for _ in range(6000):
            avg = pool[rng.integers(len(pool))]
            mn = pool[rng.integers(len(pool))]
            assert hq_loop(avg, mn, pct) == hq_single(avg, mn, pct), (e, avg, mn)
            checked += 1
    assert checked == 66000


def test_plants_are_caught():
    """A plausible slip, testing the thresholds of pct instead of hqPercent, disagrees with
    the loop on supports between the two."""
    pct = 6
    a_pct, m_pct = thresholds(pct)
    a_half, m_half = thresholds(3)
    avg = f32((a_pct + a_half) / 2)
    mn = f32(1.0)
    wrong = fgeq(avg, a_pct) and fgeq(mn, m_pct)
    assert hq_loop(avg, mn, pct) != wrong
    assert hq_loop(avg, mn, pct) == hq_single(avg, mn, pct)
