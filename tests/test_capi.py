"""CPU checks of the C ABI: the library is built, loads, and exports every
function declared in include/hundun_flow.h (no compute without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hundun_flow.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_abi():
    import hundun_flow as H
    assert declared_functions() == sorted(H.ABI_FUNCTIONS)


def test_library_exports_every_symbol():
    import hundun_flow as H
    lib = H.load_library()
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    import hundun_flow as H
    # sizes follow the C declarations (natural alignment, x86-64)
    assert ctypes.sizeof(H.hf_status) == 8
    assert ctypes.sizeof(H.hf_policy) == 40
    assert ctypes.sizeof(H.hf_fgmres_control) == 32
    assert ctypes.sizeof(H.hf_solve_result) == 96
    assert ctypes.sizeof(H.hf_solve_control) == 32
    assert ctypes.sizeof(H.hf_grid_desc) == 64
    assert ctypes.sizeof(H.hf_level_info) == 16


def test_policy_defaults_match_reference():
    import hundun_flow as H
    p = H.default_policy()
    # MgHierarchyPolicy defaults, include/hundun/v04_linear.hpp:115-131
    assert p.anisotropy_threshold == 4.0
    assert (p.pre_sweeps, p.post_sweeps, p.maximum_levels) == (2, 2, 32)
    assert (p.coarse_sweeps, p.minimum_coarse_extent) == (24, 3)
    assert p.f_cycle == 0
    # point_smoother red_black, chebyshev_lower_spectrum_fraction 0.3,
    # diagonal_shortcut_maximum_ratio 0 (off)
    assert p.point_smoother == 0
    assert p.chebyshev_lower_fraction == 0.3
    assert p.diagonal_shortcut_ratio == 0.0


def test_invalid_arguments_are_rejected_without_gpu():
    import hundun_flow as H
    lib = H.load_library()
    # null plan pointers fail with invalid_plan before any device work
    st = lib.hf_operator_apply(None, None, None)
    assert st.code == 2
    st = lib.hf_plan_create(None, None, None, None)
    assert st.code == 2
    assert lib.hf_status_message(H.hf_status(5, 0)) == b"numerical failure"


@pytest.mark.parametrize("n", [(2048, 1024, 1025), (46341, 46341, 1),
                               (2147483647, 1, 1)])
def test_grid_beyond_32bit_indexing_is_refused(n):
    """A grid whose cell or face count exceeds 2^31 - 1 is refused with
    invalid_plan before any allocation (the kernels index with int32)."""
    import numpy as np
    import hundun_flow as H
    lib = H.load_library()
    nx, ny, nz = n
    # face coordinates are only read after the size check; keep them short
    # for the huge extents by pointing at a small valid array
    f = [np.linspace(0.0, 1.0, min(v, 4096) + 1) for v in n]
    desc = H.hf_grid_desc()
    desc.nx, desc.ny, desc.nz = nx, ny, nz
    keep = [np.ascontiguousarray(v) for v in f]
    desc.faces_x = keep[0].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    desc.faces_y = keep[1].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    desc.faces_z = keep[2].ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for i, b in enumerate([1, 1, 1, 1, 1, 1]):
        desc.boundary[i] = b
    pol = H.default_policy()
    out = ctypes.c_void_p()
    st = lib.hf_plan_create(ctypes.byref(desc), ctypes.byref(pol), None,
                            ctypes.byref(out))
    assert st.code == 2 and st.detail == 9101
    assert not out.value
