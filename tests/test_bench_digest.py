"""The bench's full-size parity check (CPU only): bench.records_digest of the records of the
first n anchors, ordered by read id, is the same for the reference compiled from source
(records of anchors [0, n), sorted by read id), the restatement on the same anchors
(anchor order) and the restatement's whole run filtered to reads below n, and it changes
when one record byte changes."""
import pytest

import bench
import oracle
from tests import helpers


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_records_digest_agrees_across_runs():
    s, rs, o = helpers.case("se_q")
    n = 2000
    ref = oracle.CareRef(rs, threads=8, **o)
    _, d_ref = ref.correct(8, n)
    ref.close()
    cpu = oracle.CareCpu(rs, threads=8, **o)
    _, d_part = cpu.correct(0, n, threads=8)
    _, d_all = cpu.correct(threads=8)
    cpu.close()
    a = bench.records_digest(d_ref, n)
    assert a[0] > 100
    assert a == bench.records_digest(d_part, n) == bench.records_digest(d_all, n)
    flipped = bytearray(d_all)
    flipped[12] ^= 1                     # first payload byte of the first record (a read below n)
    assert bench.records_digest(bytes(flipped), n) != a
