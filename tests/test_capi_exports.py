"""The C-ABI library loads without a GPU and exports every entry point that
include/vpfp.h declares (no compute call is made)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("vpfp.h", "vpfc.h")]
LIB = os.path.join(ROOT, "vlapy_amd", "libvpfp.so")


def declared():
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(vpf[pc]_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_header_declares_the_boundary():
    names = declared()
    for must in ("vpfp_plan_create", "vpfp_plan_destroy", "vpfp_vdfdx", "vpfp_edfdv", "vpfp_field",
                 "vpfp_solve_for_field", "vpfp_compute_charges", "vpfp_collide", "vpfp_collision_diagonals",
                 "vpfp_tridiag_solve", "vpfp_step_n", "vpfp_moments", "vpfp_x_modes", "vpfp_last_error",
                 "vpfp_plan_autotune", "vpfp_batch_autotune", "vpfc_create", "vpfc_apply", "vpfc_train_step",
                 "vpfc_gradients", "vpfp_batch_set_closure"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvpfp.so not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    lib.vpfp_version.restype = ctypes.c_int
    assert lib.vpfp_version() >= 1


@pytest.mark.skipif(not os.path.exists(LIB), reason="libvpfp.so not built")
def test_python_binding_covers_the_header():
    from vlapy_amd import _lib

    assert set(declared()) <= set(_lib.SIGNATURES)
