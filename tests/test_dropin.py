"""The reference-side drop-in care::mi355x::correct (rabbitqcplus_amd/dropin/correct_mi355x.cpp):
compiled against the reference's own headers (options.hpp, cpureadstorage.hpp,
serializedobjectstorage.hpp) and linked to libcare_hip.so by oracle/Makefile.

CPU: the library loads, resolves the engine library, and a failing engine call surfaces as
the drop-in's std::runtime_error with the engine's message (here: no GPU in the process).
GPU (tests/test_gpu_dropin.py): its SerializedObjectStorage bytes equal correct_cpu's.
"""
import pytest

import oracle
from tests import helpers


@pytest.mark.skipif(not oracle.dropin_available(), reason="oracle/_ref/libcaredropin.so not built")
def test_dropin_reports_engine_errors_as_exceptions(monkeypatch):
    monkeypatch.setenv("HIP_VISIBLE_DEVICES", "")      # no device, here and on a GPU box
    s, rs, o = helpers.case("short")
    with pytest.raises(RuntimeError, match="care_hip_create failed with code 5"):
        oracle.dropin_correct(rs, device=0, **o)
