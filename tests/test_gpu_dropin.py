"""The reference-side drop-in care::mi355x::correct, compiled against the reference headers
and linked to libcare_hip.so, returns the SerializedObjectStorage that
care::cpu::correct_cpu returns (reference include/correct_cpu.hpp:17-21, the seam at
src/dispatch_care_correct_cpu.cpp:363): the same records, byte for byte, in read-id order
(correct_cpu's own order depends on its threads; STEP 3 sorts by read id,
src/dispatch_care_correct_cpu.cpp:409)."""
import pytest

import oracle
from tests import helpers

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not oracle.dropin_available(), reason="oracle/_ref/libcaredropin.so not built")]


@pytest.mark.parametrize("name", ["se_q", "pe", "varlen_q", "nbases"])
def test_dropin_storage_equals_correct_cpu(name):
    s, rs, o = helpers.case(name)
    print(f"[case {name}] reads={rs.n}", flush=True)
    n_d, d_d = oracle.dropin_correct(rs, device=0, **o)
    ref = oracle.CareRef(rs, threads=8, **o)
    n_r, d_r = ref.correct(threads=8)
    ref.close()
    assert n_d == n_r and d_d == d_r
