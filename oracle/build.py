"""Build the oracle (CPU restatement of the reference path) per case.

TEST INFRASTRUCTURE ONLY: the libraries built here are loaded by tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
``python -m oracle.build [case ...]`` writes oracle/_build/liboracle_<case>.so.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from vidyut3d_amd.build import case_source_dir, list_cases as _product_list_cases  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CASES = os.path.join(ROOT, "cases")
OUT = os.path.join(HERE, "_build")


def case_dim(case: str) -> int:
    with open(os.path.join(CASES, case, "GNUmakefile")) as f:
        return int(re.search(r"^\s*DIM\s*=\s*(\d)", f.read(), re.M).group(1))


# The CPU-baseline build (bench.py times it; the parity tests use the
# default -O2 build, whose rounding they were validated with): -O3 for the
# x86-64-v3 level (AVX2, FMA), which current Xeon and EPYC hosts support.
FAST_FLAGS = "g++ -O3 -march=x86-64-v3 -fopenmp"


def lib_path(case: str, variant: str = "") -> str:
    return os.path.join(OUT, f"liboracle_{case}{'_' + variant if variant else ''}.so")


def build_case(case: str, variant: str = "") -> str:
    os.makedirs(OUT, exist_ok=True)
    out = lib_path(case, variant)
    cdir = case_source_dir(case)
    if cdir is None:
        raise FileNotFoundError(f"case {case}: plugin headers not found")
    opt = ["-O3", "-march=x86-64-v3"] if variant == "fast" else ["-O2"]
    cmd = ["g++", *opt, "-std=c++17", "-fopenmp", "-fPIC", "-shared",
           f"-DAMREX_SPACEDIM={case_dim(case)}",
           "-I", cdir, "-I", os.path.join(ROOT, "vidyut3d_amd", "compat"),
           "-I", os.path.join(ROOT, "vidyut3d_amd", "core"),
           os.path.join(HERE, "oracle.cpp"), os.path.join(cdir, "Chemistry.cpp"),
           "-o", out + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed for {case}:\n{r.stderr[-4000:]}")
    os.replace(out + ".tmp", out)
    return out


def list_cases() -> list[str]:
    return _product_list_cases()


def build_all(cases=None) -> list[str]:
    out = [build_case(c) for c in (cases or list_cases())]
    if cases is None or "air3d" in cases:
        out.append(build_case("air3d", "fast"))
    return out


if __name__ == "__main__":
    for p in build_all(sys.argv[1:] or None):
        print("built", p)
