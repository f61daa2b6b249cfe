"""The C-ABI library loads without a GPU and exports every symbol of include/d2hip.h."""
import ctypes
import os

import numpy as np
import pytest

import delta2_amd as d2


def test_library_exports_every_declared_symbol():
    L = d2.lib()
    names = d2.exported_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_kernel_names():
    L = d2.lib()
    names = [L.d2_kernel_name(i).decode() for i in range(L.d2_num_kernels())]
    assert "soup_edge" in names and "tri_dist" in names


def test_null_context_is_rejected():
    L = d2.lib()
    assert L.d2_set_profiling(None, 1) == d2.D2_ERR_ARG
    assert L.d2_sync(None) == d2.D2_ERR_ARG


def test_tree_build_rejects_bad_faces():
    V = np.zeros((3, 3))
    with pytest.raises(d2.D2Error):
        d2.tree_build(V, [[0, 1, 5]], d2.pcg_seed())


def test_tree_build_small_output_keeps_rng():
    from delta2_amd import scenes
    V, F = scenes.icosphere(1)
    rng = d2.pcg_seed()
    before = rng.copy()
    n = ctypes.c_int(0)
    rc = d2.lib().d2_tree_build(d2._dptr(np.ascontiguousarray(V)), V.shape[0], d2._iptr(np.ascontiguousarray(F)),
                                F.shape[0], rng.ctypes.data_as(d2._u64p), 1, None, None, ctypes.byref(n))
    assert rc == d2.D2_ERR_CAPACITY and n.value > 1
    assert (rng == before).all()


def test_missing_library_fails_loudly(tmp_path):
    """No CPU fallback: without the built library the package refuses to run."""
    saved_lib, saved_path = d2._lib, d2.LIB_PATH
    try:
        d2._lib = None
        d2.LIB_PATH = str(tmp_path / "libd2hip.so")
        with pytest.raises(ImportError):
            d2.lib()
    finally:
        d2._lib, d2.LIB_PATH = saved_lib, saved_path


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present: contexts can be created")
def test_context_without_gpu_fails_loudly():
    """Without a GPU a context cannot be created, so no compute call can run."""
    with pytest.raises(d2.D2Error) as e:
        d2.Context(0, d2.D2_FP64)
    assert e.value.code == d2.D2_ERR_HIP
