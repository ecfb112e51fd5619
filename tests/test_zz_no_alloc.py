"""The allocation-free entry point care_mlp_forward_static (no counterpart in the
reference): through the C-ABI, its logits equal care_mlp_forward's bit for bit, on two
calls in a row, the second with fewer rows than the first; rows beyond its fixed scratch
are rejected with CARE_EINVAL and a message."""
import ctypes

import numpy as np
import pytest

from rabbitqcplus_amd import engine, learned

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    p = learned.init_params(seed)
    p[learned.OFF_B1:learned.OFF_W2] = rng.normal(0, 0.1, 64)
    p[learned.OFF_B2:learned.OFF_W3] = rng.normal(0, 0.1, 64)
    return p


def test_forward_static_equals_forward():
    lib = engine.load_library()
    h = ctypes.c_void_p()
    assert lib.care_mlp_create(0, ctypes.byref(h)) == 0
    try:
        p = np.ascontiguousarray(_params(3), np.float32)
        assert lib.care_mlp_set_params(h, p.ctypes.data, p.size) == 0
        rng = np.random.default_rng(9)
        for n in (65536, 1000):                      # the second call is smaller
            X = np.ascontiguousarray(rng.normal(0, 1, (n, 16)), np.float32)
            z_ref = np.zeros(n, np.float32)
            z_st = np.full(n, np.nan, np.float32)
            assert lib.care_mlp_forward(h, X.ctypes.data, n, z_ref.ctypes.data) == 0
            rc = lib.care_mlp_forward_static(h, X.ctypes.data, n, z_st.ctypes.data)
            assert rc == 0, lib.care_mlp_last_error(h)
            assert np.array_equal(z_ref.view(np.uint32), z_st.view(np.uint32)), n
        big = np.zeros((65537, 16), np.float32)
        out = np.zeros(65537, np.float32)
        assert lib.care_mlp_forward_static(h, big.ctypes.data, 65537, out.ctypes.data) == 1
        assert b"65536" in lib.care_mlp_last_error(h)
    finally:
        lib.care_mlp_destroy(h)
