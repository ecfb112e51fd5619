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


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_reference_with_loaded_tables_gives_the_same_records(tmp_path):
    """The CPU baseline's reference reads its tables with its own loader from a table file
    (the format the engine writes): the records equal those of the reference's own build."""
    s, rs, o = helpers.case("se_q")
    ref = oracle.CareRef(rs, threads=8, **o)
    path = str(tmp_path / "tables.bin")
    assert ref.lib.ref_save_tables(ref.h, path.encode()) == 0
    want = ref.correct(8, 5000)
    ref.close()
    ref = oracle.CareRef(rs, threads=8, tables=path, **o)
    assert ref.load_seconds is not None
    assert ref.correct(8, 5000) == want
    ref.close()
