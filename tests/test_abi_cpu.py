"""C-ABI checks that need no GPU: every case library loads, exports every
entry point declared in include/vidyut3d_amd.h, the Python host module binds
exactly that set, and a context cannot be created without a device (the
product path fails loudly; there is no CPU fallback)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vidyut3d_amd.h")
CASES = ("argon2d", "air3d", "argonrf2d")


def declared_symbols() -> set[str]:
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    src = re.sub(r"//[^\n]*", "", src)
    return set(re.findall(r"\b(vd_[a-z0-9_]+)\s*\(", src))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("vd_create", "vd_destroy", "vd_step", "vd_solve_potential", "vd_explicit_source",
              "vd_implicit_solve", "vd_reaction_source", "vd_time_scales"):
        assert s in syms


def test_python_binds_declared_set():
    from vidyut3d_amd import EXPORTED_SYMBOLS
    assert set(EXPORTED_SYMBOLS) == declared_symbols()


@pytest.mark.parametrize("case", CASES)
def test_library_exports_every_symbol(case):
    from vidyut3d_amd import lib_path
    path = lib_path(case)
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")
    lib = ctypes.CDLL(path)
    missing = [s for s in sorted(declared_symbols()) if not hasattr(lib, s)]
    assert not missing, missing
    lib.vd_kernel_class_count.restype = ctypes.c_int
    lib.vd_kernel_class_name.restype = ctypes.c_char_p
    n = lib.vd_kernel_class_count()
    names = [lib.vd_kernel_class_name(i).decode() for i in range(n)]
    assert "flux" in names and "pois_gsrb_l0" in names and len(set(names)) == n


@pytest.mark.parametrize("case", CASES)
def test_create_fails_loudly_without_gpu(case):
    import torch
    from vidyut3d_amd import Vidyut, VidyutError, lib_path
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    if not os.path.exists(lib_path(case)):
        pytest.skip("library not built")
    with pytest.raises(VidyutError) as e:
        Vidyut(case)
    assert "VD_E_HIP" in str(e.value)


def _kv(text: str) -> dict:
    out = {}
    for ln in text.splitlines():
        ln = ln.split("#", 1)[0]
        if "=" in ln:
            k, v = ln.split("=", 1)
            out[k.strip()] = v.strip()
    return out


def test_inputs_reader_overrides():
    from vidyut3d_amd import read_inputs
    p = _kv(read_inputs("air3d", {"amr.n_cell": "32 24 40", "vidyut.new_key": 7}))
    assert p["amr.n_cell"] == "32 24 40"
    assert p["vidyut.hyp_order"] == "5"
    assert p["vidyut.new_key"] == "7"
    assert _kv(read_inputs("argon2d"))["vidyut.hyp_order"] == "1"


# Options outside the path and grids the multigrid cannot handle are
# rejected while the inputs are parsed, before any device call, with
# VD_E_BADARG and a message naming the cause (the reference aborts on
# unknown options: Source/Vidyut.cpp:143,460,491,527).
REJECTED = [
    ("argon2d", {"vidyut.using_ib": 1}, "using_ib"),
    # photoionization needs the species that takes the photoelectron partner
    ("argon2d", {"vidyut.do_photoionization": 1}, "photoion_ID"),
    # RZ is two-dimensional and single-level
    ("air3d", {"geometry.coord_sys": 1}, "axisymmetric"),
    ("argon2d", {"geometry.coord_sys": 2}, "axisymmetric"),
    ("lmea3d", {"geometry.coord_sys": 1}, "axisymmetric"),
    ("argon2d", {"geometry.coord_sys": 1, "amr.max_level": 1}, "Cartesian"),
    ("argon2d", {"vidyut.adaptive_dt": 1}, "adaptive_dt"),
    ("argon2d", {"vidyut.hyp_order": 3}, "hyp_order"),
    ("air3d", {"vidyut.weno_scheme": 4}, "WENO"),
    ("argon2d", {"amr.n_cell": "32 15"}, "even number of cells"),     # y periodic, odd extent
    ("argon2d", {"amr.n_cell": "250 16"}, "coarsest multigrid level"),  # 250 -> 125: odd, > 32
    ("argon2d", {"amr.n_cell": "64 64", "vidyut.linsolve_max_coarsening_level": 0},
     "coarsest multigrid level"),
    ("argon2d", {"amr.max_level": 1, "amr.n_cell": "34 16"}, "divisible by 4"),
    ("argon2d", {"amr.max_level": 1, "vidyut.num_timestep_correctors": 2}, "num_timestep_correctors"),
    ("argonrf2d", {"vidyut.current_loc_surfaces": "1 7"}, "current_loc_surfaces"),
]


@pytest.mark.parametrize("case,ov,why", REJECTED)
def test_create_rejects_unsupported_inputs(case, ov, why):
    from vidyut3d_amd import VD_OK, lib_path, load_library, read_inputs
    if not os.path.exists(lib_path(case)):
        pytest.skip("library not built")
    lib = load_library(case)
    h = ctypes.c_void_p()
    rc = lib.vd_create(read_inputs(case, ov).encode(), ctypes.byref(h))
    assert rc == 1, (rc, lib.vd_global_error())   # VD_E_BADARG
    assert not h.value
    assert why in lib.vd_global_error().decode()
    assert rc != VD_OK
