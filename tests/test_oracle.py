"""Pins the CPU oracle (oracle/carecpu.cpp).

Checks, per part of the path:
  * golden fixtures written by the reference engine compiled from source
    (tests/golden/make_golden.py): records of 10 workloads, signatures, candidate
    lists, alignments, orientation decisions, MSA columns;
  * brute force from the definitions (k-mer enumeration + MurmurHash3 fmix64,
    dictionary grouping for the tables, exhaustive shift scan for the alignment),
    written independently of the oracle's code;
  * the live reference engine on fresh seeds where oracle/_ref is built.
"""
import os

import numpy as np
import pytest

import oracle
from tests import helpers
from tests.golden import make_golden as mg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MASK64 = (1 << 64) - 1


def load(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


# ------------------------------------------------------------------ records

@pytest.mark.parametrize("name", mg.RECORD_CASES)
def test_oracle_records_match_reference_golden(name):
    s, rs, o = helpers.case(name)
    fx = load(f"records_{name}.npz")
    assert mg.storage_digest(rs) == str(fx["digest"]), "simulator output changed; regenerate fixtures"
    cpu = oracle.CareCpu(rs, threads=8, **o)
    n, data = cpu.correct(threads=8)
    cpu.close()
    assert n == int(fx["num_records"])
    assert data == fx["records"].tobytes()


# ------------------------------------------------------------------ signatures

def fmix64(x):
    x ^= x >> 33
    x = (x * 0xff51afd7ed558ccd) & MASK64
    x ^= x >> 33
    x = (x * 0xc4ceb9fe1a85ec53) & MASK64
    x ^= x >> 33
    return x


def brute_signature(seq: bytes, k: int, nmaps: int):
    code = {ord("A"): 0, ord("C"): 1, ord("G"): 2, ord("T"): 3}
    mask = (1 << (2 * k)) - 1
    if len(seq) < k:
        return [mask] * nmaps
    mins = [MASK64] * nmaps
    for p in range(len(seq) - k + 1):
        f = 0
        r = 0
        for i in range(k):
            f = (f << 2) | code[seq[p + i]]
            r = (r << 2) | (3 - code[seq[p + k - 1 - i]])
        c = min(f, r)
        for m in range(nmaps):
            mins[m] = min(mins[m], fmix64((c + m) & MASK64))
    return [v & mask for v in mins]


def test_signatures_golden():
    fx = load("stages_se.npz")
    s, rs, o = helpers.case("se")
    assert mg.storage_digest(rs) == str(fx["digest"])
    for key, k, m in (("sig20", 20, 48), ("sig32", 32, 48), ("sig11", 11, 16)):
        got = np.stack([oracle.cpu_hash(rs.enc[i], rs.lengths[i], k, m) for i in fx["sig_ids"]])
        assert np.array_equal(got, fx[key]), key


def test_signatures_bruteforce():
    s, rs, o = helpers.case("short")
    rng = np.random.default_rng(1)
    for i in rng.choice(rs.n, 6, replace=False):
        seq = rs.decoded(int(i))
        for k, m in ((20, 6), (8, 3)):
            assert list(oracle.cpu_hash(rs.enc[i], rs.lengths[i], k, m)) == brute_signature(seq, k, m)


# ------------------------------------------------------------------ tables / candidates

def test_candidates_golden():
    fx = load("stages_se.npz")
    s, rs, o = helpers.case("se")
    cpu = oracle.CareCpu(rs, threads=8, **o)
    offs = fx["cand_offsets"]
    for j, a in enumerate(fx["cand_anchors"]):
        assert np.array_equal(cpu.candidates(int(a)), fx["cand_ids"][offs[j]:offs[j + 1]])
    cpu.close()


@pytest.mark.parametrize("name", ["short", "se_repeats"])
def test_candidates_bruteforce(name):
    """Candidates = union over maps of the reads sharing the anchor's signature in
    buckets of 2..threshold reads, minus the anchor (reads shorter than k query nothing)."""
    s, rs, o = helpers.case(name)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    sig = cpu.signatures()
    thr = cpu.lib.carecpu_results_per_map_threshold(float(o.get("estimated_coverage", 30.0)))
    k = o.get("kmerlength", 20)
    buckets = []
    for m in range(sig.shape[1]):
        d = {}
        for r, key in enumerate(sig[:, m].tolist()):
            d.setdefault(key, []).append(r)
        buckets.append(d)
    rng = np.random.default_rng(3)
    for a in rng.choice(rs.n, 200, replace=False):
        a = int(a)
        want = set()
        if rs.lengths[a] >= k:
            for m in range(sig.shape[1]):
                b = buckets[m][int(sig[a, m])]
                if 2 <= len(b) <= thr:
                    want.update(b)
        want.discard(a)
        assert cpu.candidates(a).tolist() == sorted(want)
    cpu.close()


def test_threshold_rule():
    lib = oracle.cpu_lib()
    # max(10, int(int(coverage) * 2.5f)), capped at 65535
    for cov, want in ((30.0, 75), (1.0, 10), (4.9, 10), (30.9, 75), (31.0, 77), (100000.0, 65535)):
        assert lib.carecpu_results_per_map_threshold(cov) == want


# ------------------------------------------------------------------ alignment

def brute_align(a: bytes, c: bytes, min_overlap=30, max_error_rate=0.2, ratio=0.3):
    """Exhaustive scan: admissible shifts have mismatches < int(float32(overlap)*0.2f);
    the winner is the first admissible shift, in the order 0..La-mo, -1..-(Lc-mo), of
    minimum mismatches + La + Lc - 2*overlap."""
    A = np.frombuffer(a, np.uint8)
    C = np.frombuffer(c, np.uint8)
    La, Lc = len(A), len(C)
    total = La + Lc
    mo = max(min_overlap, int(np.float32(La) * np.float32(ratio)))
    order = list(range(0, La - mo + 1)) + [-s for s in range(1, Lc - mo + 1)]
    best, best_shift = total, -Lc
    for s in order:
        ov = min(La - s, Lc) if s >= 0 else min(La, Lc + s)
        if ov <= 0:
            continue
        lim = int(np.float32(ov) * np.float32(max_error_rate))
        mism = int((A[s:s + ov] != C[:ov]).sum()) if s >= 0 else int((C[-s:-s + ov] != A[:ov]).sum())
        if mism < lim:
            score = mism + total - 2 * ov
            if score < best:
                best, best_shift = score, s
    cb = max(-best_shift, 0)
    ce = min(Lc, La - best_shift)
    ov = ce - cb
    return [best, ov, best_shift, best - total + 2 * ov, int(best_shift != -Lc)]


def test_alignment_golden():
    fx = load("stages_se.npz")
    stores = [helpers.case("se")[1], helpers.case("varlen_q")[1]]
    assert mg.storage_digest(stores[1]) == str(fx["digest_varlen"])
    for (st, a, c), want in zip(fx["align_pairs"], fx["align_results"]):
        R = stores[st]
        got = oracle.cpu_align(R.enc[a], int(R.lengths[a]), R.enc[c], int(R.lengths[c]))
        assert list(got) == list(want), (st, a, c)


def test_alignment_bruteforce():
    rng = np.random.default_rng(5)
    s, rs, o = helpers.case("se_higherr")
    cpu = oracle.CareCpu(rs, threads=8, **o)
    pairs = []
    for a in rng.choice(rs.n, 12, replace=False):
        for c in cpu.candidates(int(a))[:8]:
            pairs.append((int(a), int(c)))
    cpu.close()
    for _ in range(40):
        pairs.append((int(rng.integers(rs.n)), int(rng.integers(rs.n))))
    comp = bytes.maketrans(b"ACGT", b"TGCA")
    for a, c in pairs:
        A, Cs = rs.decoded(a), rs.decoded(c)
        for cand, row in ((Cs, rs.enc[c]), (Cs.translate(comp)[::-1], None)):
            if row is None:
                from rabbitqcplus_amd import reads as R
                codes = np.frombuffer(cand, np.uint8)
                lut = np.zeros(256, np.uint8)
                lut[[65, 67, 71, 84]] = [0, 1, 2, 3]
                row = R.encode_codes(lut[codes][None, :], np.array([len(cand)], np.int32))[0]
            got = oracle.cpu_align(rs.enc[a], len(A), row, len(cand))
            assert list(got) == brute_align(A, cand), (a, c)


def test_alignment_adversarial_golden():
    fx = load("stages_se.npz")
    for seqs, lens, want in zip(fx["adv_seq"], fx["adv_len"], fx["adv_res"]):
        a = seqs[0, :lens[0]].tobytes()
        c = seqs[1, :lens[1]].tobytes()
        got = oracle.cpu_align(mg.encode(a), len(a), mg.encode(c), len(c))
        assert list(got) == list(want)
        assert brute_align(a, c) == list(want)


def test_alignment_planted_tie_break_is_detected():
    """A plausible mistake (last instead of first minimum in scan order) changes
    results on the golden low-complexity pairs, so the golden set pins the tie-break."""
    fx = load("stages_se.npz")
    differs = 0
    for seqs, lens, want in zip(fx["adv_seq"], fx["adv_len"], fx["adv_res"]):
        A = seqs[0, :lens[0]].tobytes()
        C = seqs[1, :lens[1]].tobytes()
        La, Lc = len(A), len(C)
        Aa, Cc = np.frombuffer(A, np.uint8), np.frombuffer(C, np.uint8)
        mo = max(30, int(np.float32(La) * np.float32(0.3)))
        order = list(range(0, La - mo + 1)) + [-s for s in range(1, Lc - mo + 1)]
        best, bs = La + Lc, -Lc
        for sft in order:
            ov = min(La - sft, Lc) if sft >= 0 else min(La, Lc + sft)
            lim = int(np.float32(ov) * np.float32(0.2))
            mism = int((Aa[sft:sft + ov] != Cc[:ov]).sum()) if sft >= 0 else int((Cc[-sft:-sft + ov] != Aa[:ov]).sum())
            if mism < lim and mism + La + Lc - 2 * ov <= best:     # planted: '<=' keeps the last minimum
                best, bs = mism + La + Lc - 2 * ov, sft
        differs += int(bs != want[2])
    assert differs > 0


def test_orientation_golden():
    fx = load("stages_se.npz")
    lib = oracle.cpu_lib()
    for (fo, ro, fn, rn, fv, rv, la, lc), want in zip(fx["orient_in"], fx["orient_out"]):
        f5 = np.array([0, fo, 0, fn, fv], np.int32)
        r5 = np.array([0, ro, 0, rn, rv], np.int32)
        got = lib.carecpu_orientation(f5.ctypes.data, r5.ctypes.data, int(la), int(lc), 0.3, 30, 0.06)
        assert got == want


# ------------------------------------------------------------------ MSA

@pytest.mark.parametrize("name", ["se_higherr", "se_higherr_q"])
def test_msa_columns_golden(name):
    fx = load("msa_columns.npz")
    s, rs, o = helpers.case(name)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    off = 0
    for a, ncol, beg in zip(fx[f"{name}_anchors"], fx[f"{name}_ncol"], fx[f"{name}_begin"]):
        tr = cpu.trace(int(a))
        sl = slice(off, off + int(ncol))
        off += int(ncol)
        assert tr["num_columns"] == ncol and tr["anchor_column_begin"] == beg
        assert np.array_equal(tr["counts"], fx[f"{name}_counts"][sl])
        assert np.array_equal(tr["coverage"], fx[f"{name}_coverage"][sl])
        assert np.array_equal(tr["consensus"], fx[f"{name}_consensus"][sl])
        assert np.array_equal(tr["weights"], fx[f"{name}_weights"][sl])
        assert np.array_equal(tr["support"], fx[f"{name}_support"][sl])
        # invariants: coverage = sum of counts; consensus = first argmax of the weights
        assert np.array_equal(tr["coverage"], tr["counts"].sum(axis=1))
        assert np.array_equal(np.frombuffer(b"ACGT", np.uint8)[tr["weights"].argmax(axis=1)], tr["consensus"])
    cpu.close()


# ------------------------------------------------------------------ live reference

@pytest.mark.skipif(not oracle.ref_available(), reason="reference engine not built (make -C oracle ref)")
@pytest.mark.parametrize("seed,paired,q", [(101, 0, 0), (102, 0, 1), (103, 1, 1)])
def test_oracle_vs_reference_fresh_seeds(seed, paired, q):
    from rabbitqcplus_amd import reads, sim
    sr = sim.simulate(genome_length=40_000, read_length_min=120, read_length_max=150, coverage=30.0,
                      q_first=26.0, q_last=16.0, paired=paired, seed=seed, n_rate=0.0005)
    rs = reads.storage_from_sim(sr, use_qualities=bool(q))
    o = dict(use_quality_scores=q, paired=paired, kmerlength=20)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    n1, d1 = cpu.correct(threads=8)
    cpu.close()
    ref = oracle.CareRef(rs, threads=8, **o)
    n2, d2 = ref.correct(threads=8)
    ref.close()
    assert n1 == n2 and d1 == d2


def _pair_lists(rng, universe, n1, n2):
    c1 = np.unique(rng.choice(universe, n1)).astype(np.uint32)
    c2 = np.unique(rng.choice(universe, n2)).astype(np.uint32)
    return c1, c2


def test_pair_flags_equal_membership_rule():
    """The GPU sets paired flags by membership: x in one mate's list is flagged exactly
    when x ^ 1 (the other read of x's pair) is in the other mate's list. The reference's
    merge state machine (include/util.hpp:1016-1141) must give the same flags on dense,
    sparse, empty and overlapping sorted lists."""
    rng = np.random.default_rng(5)
    cases = [(np.array([], np.uint32), np.array([], np.uint32)),
             (np.array([4, 5], np.uint32), np.array([], np.uint32)),
             (np.array([4, 5], np.uint32), np.array([4, 5], np.uint32)),
             (np.array([4], np.uint32), np.array([4], np.uint32)),
             (np.array([4, 5], np.uint32), np.array([4], np.uint32)),
             (np.array([4, 5], np.uint32), np.array([5], np.uint32)),
             (np.array([5], np.uint32), np.array([4, 5, 6], np.uint32))]
    for universe, n in [(16, 10), (64, 30), (400, 120), (4000, 300), (1 << 20, 500)]:
        for _ in range(60):
            cases.append(_pair_lists(rng, universe, int(rng.integers(0, n)), int(rng.integers(0, n))))
    flagged = 0
    for c1, c2 in cases:
        m1 = np.isin(c1 ^ np.uint32(1), c2)
        m2 = np.isin(c2 ^ np.uint32(1), c1)
        f1, f2 = oracle.cpu_flag_pairs(c1, c2)
        assert np.array_equal(f1, m1) and np.array_equal(f2, m2), (c1, c2)
        if oracle.ref_available():
            r1, r2 = oracle.ref_flag_pairs(c1, c2)
            assert np.array_equal(r1, m1) and np.array_equal(r2, m2), (c1, c2)
        flagged += int(m1.sum())
    assert flagged > 500
