"""The parity cases reach every size-dependent route of the HIP path (CPU only, counted
with the oracle's candidate lists): anchors with at most and with more than 64 candidates
(the one-wavefront MSA keeps lists of at most 64 in lanes and longer ones in memory; the
lane alignment takes longer lists in several 64-lane chunks) and anchors with more than
256 unique candidates (the candidate sort's bitonic route after the rank sort)."""
import numpy as np
import pytest

import oracle
from tests import helpers


def candidate_counts(name):
    s, rs, o = helpers.case(name)
    cpu = oracle.CareCpu(rs, threads=8, **o)
    try:
        return np.array([len(cpu.candidates(i)) for i in range(rs.n)])
    finally:
        cpu.close()


@pytest.mark.parametrize("name, above64, above256", [("se", True, False), ("se_repeats", True, True)])
def test_cases_cover_list_size_routes(name, above64, above256):
    cnt = candidate_counts(name)
    assert (cnt <= 64).sum() > 1000
    assert ((cnt > 64).sum() > 100) == above64
    assert ((cnt > 256).sum() > 0) == above256
