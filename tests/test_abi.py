"""The C ABI library loads and exports every entry point include/qvort_gpu.h
declares; argument checks work without a GPU (no compute call is made)."""
import ctypes as C
import os
import re

import pytest

from conftest import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "qvort_gpu.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(qv_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    from qvort_amd import gpu
    lib = gpu.lib()
    names = _declared()
    assert "qv_velocity" in names and "qv_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(gpu.EXPORTS) == set(names)


def test_param_validation_before_device():
    from qvort_amd import gpu
    lib = gpu.lib()
    h = C.c_void_p()
    bad = [gpu.params(box_size=0.0), gpu.params(velocity="Tree", tree_theta=0.0),
           gpu.params(deriv_order=3)]
    for p in bad:
        assert lib.qv_create(C.byref(p), C.byref(h)) == gpu.QV_ERR_ARG
    p = gpu.params(velocity="Tree", tree_theta=0.4)
    p.precision = 7
    assert lib.qv_create(C.byref(p), C.byref(h)) == gpu.QV_ERR_ARG


def test_status_strings():
    from qvort_amd import gpu
    s = gpu.status_string(gpu.QV_ERR_SINGULAR)
    assert "singularity in BS (tree) velocity field" in s
    assert gpu.status_string(gpu.QV_OK) == "ok"


def test_no_cpu_fallback_without_device():
    """Without a visible GPU the library refuses to run (no CPU path)."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from qvort_amd import gpu
    with pytest.raises(gpu.QvortError) as ei:
        gpu.Context(gpu.params(velocity="Tree", tree_theta=0.4))
    assert ei.value.status == gpu.QV_ERR_NO_DEVICE


def test_batch_argument_checks_before_device():
    """Ensemble contexts take velocity='BS' only and systems of >= 3 slots;
    arguments are checked before any device call."""
    import numpy as np
    import torch
    from qvort_amd import gpu
    lib = gpu.lib()
    h = C.c_void_p()
    counts = np.array([200, 5000, 3], dtype=np.int32)
    cp = counts.ctypes.data_as(C.POINTER(C.c_int32))
    tree = gpu.params(velocity="Tree", tree_theta=0.4)
    assert lib.qv_batch_create(C.byref(tree), 3, cp, C.byref(h)) == gpu.QV_ERR_ARG
    bs = gpu.params(velocity="BS", boundary="open")
    assert lib.qv_batch_create(C.byref(bs), 0, cp, C.byref(h)) == gpu.QV_ERR_ARG
    short = np.array([200, 2], dtype=np.int32)
    assert lib.qv_batch_create(C.byref(bs), 2, short.ctypes.data_as(C.POINTER(C.c_int32)),
                               C.byref(h)) == gpu.QV_ERR_ARG
    assert not h.value
    if not torch.cuda.is_available():
        assert lib.qv_batch_create(C.byref(bs), 3, cp, C.byref(h)) == gpu.QV_ERR_NO_DEVICE
        assert lib.qv_batch_destroy(None) == gpu.QV_OK
