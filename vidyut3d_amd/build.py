"""Build the per-case gfx950 libraries of vidyut3d_amd.

One shared library per case directory (``libvd_<case>.so``), mirroring the
reference's one executable per case: the case plugin headers are compiled
into the kernels.  ``python -m vidyut3d_amd.build [case ...]``.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CASES = os.path.join(ROOT, "cases")
LIBDIR = os.path.join(PKG, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"


REFERENCE_ROOT = os.environ.get("VD_REFERENCE_ROOT", "/root/reference")


def case_source_dir(case: str) -> str | None:
    """Directory holding a case's plugin headers: ``cases/<case>`` itself, or
    the reference-tree directory named by ``cases/<case>/CASE_SOURCE`` (a
    reference case compiled unchanged where it lies).  None when that
    directory is not present on this machine."""
    cdir = os.path.join(CASES, case)
    if os.path.isfile(os.path.join(cdir, "Chemistry.H")):
        return cdir
    src = os.path.join(cdir, "CASE_SOURCE")
    if os.path.isfile(src):
        with open(src) as f:
            rel = [ln.strip() for ln in f if ln.strip() and not ln.lstrip().startswith("#")][0]
        d = os.path.join(REFERENCE_ROOT, rel)
        return d if os.path.isfile(os.path.join(d, "Chemistry.H")) else None
    return None


def case_dim(case_dir: str) -> int:
    with open(os.path.join(case_dir, "GNUmakefile")) as f:
        m = re.search(r"^\s*DIM\s*=\s*(\d)", f.read(), re.M)
    if not m:
        raise RuntimeError(f"{case_dir}/GNUmakefile has no DIM")
    return int(m.group(1))


def list_cases() -> list[str]:
    """Cases whose plugin headers are available here."""
    return sorted(d for d in os.listdir(CASES) if case_source_dir(d) is not None)


def lib_path(case: str) -> str:
    return os.path.join(LIBDIR, f"libvd_{case}.so")


def build_case(case: str, verbose: bool = False) -> str:
    cdir = case_source_dir(case)
    if cdir is None:
        raise FileNotFoundError(f"case {case}: plugin headers not found")
    dim = case_dim(os.path.join(CASES, case))
    os.makedirs(LIBDIR, exist_ok=True)
    out = lib_path(case)
    tmp = out + ".tmp"
    cmd = [
        HIPCC, f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-DAMREX_SPACEDIM={dim}",
        "-I", cdir,
        "-I", os.path.join(PKG, "compat"),
        "-I", os.path.join(PKG, "core"),
        "-I", os.path.join(PKG, "csrc"),
        os.path.join(PKG, "csrc", "vd_lib.hip"),
        "-x", "hip", os.path.join(cdir, "Chemistry.cpp"),
        "-o", tmp,
        "-L/opt/rocm/lib", "-lrocblas",
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for case {case}:\n{r.stderr[-6000:]}")
    os.replace(tmp, out)
    return out


def build_all(cases: list[str] | None = None, verbose: bool = False) -> list[str]:
    cases = cases or list_cases()
    with ThreadPoolExecutor(max_workers=min(4, len(cases))) as ex:
        return list(ex.map(lambda c: build_case(c, verbose), cases))


if __name__ == "__main__":
    for p in build_all(sys.argv[1:] or None, verbose=True):
        print("built", p)
