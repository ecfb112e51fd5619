"""Table files in the reference's minhasher format (GPU).

Bar: the file the engine writes equals the reference's own file for the same reads, field
by field (the two padding bytes of each 16-byte slot are not part of the format); tables
the reference wrote load into the engine and give the reference's records; tables the
engine wrote load into the reference (its own loader) and give its records.
"""
import os
import struct

import numpy as np
import pytest

import oracle
from rabbitqcplus_amd import engine
from tests import helpers

pytestmark = pytest.mark.gpu

CASES = ["se_q", "pe", "varlen_q"]


def parse_table_file(path):
    """Header and per-map (values, slots) of a minhasher file
    (include/ordinaryminhasher.hpp:303-339, include/cpuhashtable.hpp:221-248,627-649)."""
    blob = open(path, "rb").read()
    k, thr, lf, maps = struct.unpack_from("<iifi", blob, 0)
    p = 16
    out = []
    for _ in range(maps):
        (elements,) = struct.unpack_from("<Q", blob, p); p += 8
        values = np.frombuffer(blob, "<u4", elements, p).copy(); p += 4 * elements
        load, num_keys, max_probes, size, capacity, slots = struct.unpack_from("<fQQQQQ", blob, p); p += 44
        raw = np.frombuffer(blob, np.uint8, 16 * slots, p).reshape(slots, 16).copy(); p += 16 * slots
        keys = raw[:, 0:8].copy().view("<u8").ravel()
        offs = raw[:, 8:12].copy().view("<u4").ravel()
        counts = raw[:, 12:14].copy().view("<u2").ravel()
        out.append(dict(values=values, load=load, num_keys=num_keys, max_probes=max_probes, size=size,
                        capacity=capacity, keys=keys, offsets=offs, counts=counts))
    assert p == len(blob)
    return (k, thr, lf, maps), out


def _engine(rs, o):
    eng = engine.CareEngine(0, **o)
    eng.set_reads(rs)
    return eng


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", CASES)
def test_saved_tables_equal_reference_file(name, tmp_path):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    eng = _engine(rs, o)
    eng.build_tables()
    ours = str(tmp_path / "ours.bin")
    eng.save_tables(ours)
    eng.close()
    ref = oracle.CareRef(rs, threads=8, **o)
    theirs = str(tmp_path / "ref.bin")
    ref.save_tables(theirs)
    ref.close()
    h1, m1 = parse_table_file(ours)
    h2, m2 = parse_table_file(theirs)
    assert h1 == h2
    for a, b in zip(m1, m2):
        for key in ("load", "num_keys", "max_probes", "size", "capacity"):
            assert a[key] == b[key], key
        for key in ("values", "keys", "offsets", "counts"):
            assert np.array_equal(a[key], b[key]), key


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("name", CASES)
def test_table_files_load_both_ways(name, tmp_path):
    s, rs, o = helpers.case(name)
    golden = np.load(os.path.join(os.path.dirname(__file__), "golden", f"records_{name}.npz"))["records"].tobytes()
    # reference file -> engine
    ref = oracle.CareRef(rs, threads=8, **o)
    theirs = str(tmp_path / "ref.bin")
    ref.save_tables(theirs)
    eng = _engine(rs, o)
    assert eng.load_tables(theirs) == 48
    assert eng.correct()[1] == golden
    # engine file -> reference loader
    ours = str(tmp_path / "ours.bin")
    eng.save_tables(ours)
    eng.close()
    assert ref.load_tables(ours) == 48
    assert ref.correct(threads=8)[1] == golden
    ref.close()


def test_load_rejects_mismatched_options(tmp_path):
    s, rs, o = helpers.case("se")
    eng = _engine(rs, o)
    eng.build_tables()
    path = str(tmp_path / "t.bin")
    eng.save_tables(path)
    eng.close()
    other = _engine(rs, dict(o, kmerlength=16))
    with pytest.raises(engine.CareError, match="k-mer length"):
        other.load_tables(path)
    other.close()
    blob = open(path, "rb").read()
    short = str(tmp_path / "short.bin")
    open(short, "wb").write(blob[: len(blob) // 2])
    eng = _engine(rs, o)
    with pytest.raises(engine.CareError, match="truncated|larger than the file"):
        eng.load_tables(short)
    eng.close()
