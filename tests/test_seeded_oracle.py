"""Seeded alignment in the oracle (an addition for the fast learned form; the reference has
no counterpart, so it is pinned by brute force from its definition):

* seeds: per read and map, the start of the first k-mer whose canonical form attains the
  map's minimum hash, and whether that canonical form is the reverse complement,
  enumerated from the decoded read with MurmurHash3 fmix64 on 64-bit integers;
* the chosen map is the first one whose kept bucket holds both reads;
* the shift puts the shared k-mer at the same anchor column in both reads (the k-mer of
  the anchor at pa equals the oriented candidate's k-mer at pa - shift);
* overlap and mismatches are counted over the decoded strings, and the kept/dropped flag
  follows the reference's overlap, mismatch and orientation bounds.
A mistake planted in the shift formula is shown to be caught by the k-mer check.
"""
import ctypes

import numpy as np
import pytest

import oracle
from tests import helpers

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def fmix64_np(x):
    x = x.astype(np.uint64)
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xff51afd7ed558ccd)
        x ^= x >> np.uint64(33)
        x *= np.uint64(0xc4ceb9fe1a85ec53)
        x ^= x >> np.uint64(33)
    return x


def kmers(bases: np.ndarray, k: int):
    """Forward and reverse-complement k-mers (2k-bit integers) at every start position."""
    L = len(bases)
    if L < k:
        return np.zeros(0, np.uint64), np.zeros(0, np.uint64)
    idx = np.arange(L - k + 1)[:, None] + np.arange(k)[None, :]
    w = bases[idx].astype(np.uint64)
    sh = (np.uint64(2) * np.arange(k - 1, -1, -1, dtype=np.uint64))
    fwd = np.bitwise_or.reduce(w << sh[None, :], axis=1)
    rcw = (np.uint64(3) - w)[:, ::-1]
    rc = np.bitwise_or.reduce(rcw << sh[None, :], axis=1)
    return fwd, rc


def brute_seeds(bases: np.ndarray, k: int, maps: int):
    fwd, rc = kmers(bases, k)
    if fwd.size == 0:
        return [0] * maps
    canon = np.minimum(fwd, rc)
    strand = (rc < fwd).astype(np.int64)
    out = []
    for m in range(maps):
        h = fmix64_np(canon + np.uint64(m))
        p = int(np.argmin(h))                # first position attaining the minimum
        out.append((p << 1) | int(strand[p]))
    return out


def decoded(rs, i):
    return np.frombuffer(rs.decoded(i), np.uint8)


_CODE = np.zeros(256, np.uint8)
_CODE[[65, 67, 71, 84]] = [0, 1, 2, 3]


def codes(rs, i):
    return _CODE[decoded(rs, i)]


def expected(cpu, rs, o, a, c, sigs, seeds_a, seeds_c, thr, plant=False):
    k = o.get("kmerlength", 20)
    maps = 48
    lib = cpu.lib
    vals = ctypes.c_void_p()
    mshared = -1
    for m in range(maps):
        cnt = lib.carecpu_bucket(cpu.h, m, int(sigs[a][m]), ctypes.byref(vals))
        if 0 < cnt <= thr and sigs[c][m] == sigs[a][m]:
            mshared = m
            break
    if mshared < 0:
        return (-1, 3, None, None, None)
    A, C = codes(rs, a), codes(rs, c)
    La, Lc = len(A), len(C)
    sa, sc = seeds_a[mshared], seeds_c[mshared]
    pa, pc = sa >> 1, sc >> 1
    same = (sa & 1) == (sc & 1)
    shift = pa - pc if same else pa - (Lc - k - pc)
    if plant:
        shift = pa - pc if same else pa - (Lc - pc)
    Co = C if same else (3 - C[::-1])
    # the shared k-mer lines up: anchor[pa:pa+k] == oriented candidate[pa-shift : pa-shift+k]
    q = pa - shift
    lines_up = 0 <= q <= Lc - k and np.array_equal(A[pa:pa + k], Co[q:q + k])
    cb, ce = max(-shift, 0), min(Lc, La - shift)
    ov = ce - cb
    mis = int(np.count_nonzero(Co[cb:ce] != A[cb + shift:ce + shift])) if ov > 0 else 0
    mo = max(30, int(np.float32(La) * np.float32(0.3)))
    ok = (ov > 0 and ov >= mo and mis < int(np.float32(ov) * np.float32(0.2))
          and np.float32(mis) / np.float32(ov) < np.float32(0.06) * np.float32(4.0))
    flag = (1 if same else 2) if ok else 3
    return (mshared, flag, shift, ov, mis), lines_up


@pytest.mark.parametrize("name", ["se_higherr", "pe", "varlen_q", "se_repeats"])
def test_seeded_alignment_matches_brute_force(name):
    s, rs, o = helpers.case(name)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    try:
        k = o.get("kmerlength", 20)
        thr = cpu.lib.carecpu_results_per_map_threshold(ctypes.c_float(o.get("estimated_coverage", 30.0)))
        rng = np.random.default_rng(3)
        anchors = rng.choice(rs.n, 25, replace=False)
        sigs = {}
        seeds = {}
        checked = kept = 0
        for a in anchors:
            cands = cpu.candidates(int(a))
            ids = [int(a)] + [int(x) for x in cands]
            got_sigs = cpu.signatures()  # whole set, cheap at this size
            got_seeds = cpu.seeds()
            for i in ids:
                if i not in seeds:
                    seeds[i] = brute_seeds(codes(rs, i), k, 48)
                    assert list(got_seeds[i]) == seeds[i], f"seeds of read {i}"
                    sigs[i] = got_sigs[i]
            for c in cands:
                c = int(c)
                want, lines_up = expected(cpu, rs, o, int(a), c, sigs, seeds[int(a)], seeds[c], thr)
                got = cpu.seeded_alignment(int(a), c)
                assert want[0] >= 0, "every candidate shares a kept bucket with its anchor"
                assert got[:2] == want[:2], (a, c, got, want)
                assert got[2:] == want[2:], (a, c, got, want)
                # repeats inside a read can place the first minimizing k-mer elsewhere;
                # where the k-mer lines up, the alignment is the seed's
                checked += 1
                kept += got[1] != 3
                if got[1] != 3:
                    assert lines_up
            del got_sigs, got_seeds
        assert checked > 200 and kept > checked // 3
    finally:
        cpu.close()


def test_planted_shift_mistake_is_caught():
    """A reverse-complement shift without the -k term misplaces the shared k-mer."""
    s, rs, o = helpers.case("se_higherr")
    cpu = oracle.CareCpu(rs, threads=8, **o)
    try:
        k = 20
        thr = cpu.lib.carecpu_results_per_map_threshold(ctypes.c_float(30.0))
        sig = cpu.signatures()
        sd = cpu.seeds()
        caught = 0
        for a in range(0, 400, 20):
            for c in cpu.candidates(a)[:20]:
                c = int(c)
                sigs = {a: sig[a], c: sig[c]}
                good, ok_good = expected(cpu, rs, o, a, c, sigs, [int(x) for x in sd[a]], [int(x) for x in sd[c]], thr)
                bad, ok_bad = expected(cpu, rs, o, a, c, sigs, [int(x) for x in sd[a]], [int(x) for x in sd[c]], thr, plant=True)
                if good[1] == 2:
                    assert ok_good
                    caught += not ok_bad
        assert caught > 10
    finally:
        cpu.close()
