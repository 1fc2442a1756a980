"""Plotfile and checkpoint I/O (vidyut3d_amd/csrc/vd_io.H; Source/Plot.cpp:
32-225) on the CPU, through two implementations that share no code: the
C++ writer and reader (driven by tests/io_driver.cpp, built here with g++)
and the Python reader vidyut3d_amd.plotfile.

* A two-level plotfile and checkpoint written by the C++ side read back by
  the Python reader: every header field (format name, variable names,
  dimension, time, finest level, problem extents, refinement ratio, level
  domains and steps, cell sizes, coordinate system, physical box extents,
  MultiFab paths; the checkpoint's title, istep, dt, t_new, BoxArrays), the
  per-box minima and maxima, and every value (exact binary values, compared
  bitwise).  The FAB header lines match the pattern of yt's BoxLib reader.
* A MultiFab of several boxes in two data files, with ghost cells and
  big-endian data, written here from the documented layout, read by the C++
  reader (the restart path for a decomposed reference checkpoint) and by
  the Python reader; a box that leaves part of a level uncovered is an
  error.

The format follows AMReX's published layout (no AMReX reader exists here),
so agreement with AMReX's own reader is unpinned; these tests pin the two
implementations against each other and against the layout."""
from __future__ import annotations

import os
import re
import subprocess

import numpy as np
import pytest

from vidyut3d_amd.plotfile import read_checkpoint, read_multifab, read_plotfile

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# yt's BoxLib frontend pattern for a FAB header line
YT_FAB = re.compile(r"^FAB \(\(\d+, \([0-9 ]+\)\),\((\d+), \(([0-9 ]+)\)\)\)"
                    r"\(\((\d+(,\d+)*)\) \((\d+(,\d+)*)\) \((\d+(,\d+)*)\)\) (\d+)\n$")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("io") / "io_driver")
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(ROOT, "vidyut3d_amd", "csrc"),
                    os.path.join(ROOT, "tests", "io_driver.cpp"), "-o", exe], check=True)
    return exe


def value(c, i, j, k, l):
    return c * 1000003.0 - i * 101.0 + j * 7.0 + k * 0.5 + l * 0.25


def expected(box_lo, n, ncomp, level):
    c, k, j, i = np.meshgrid(np.arange(ncomp), np.arange(n[2]) + box_lo[2], np.arange(n[1]) + box_lo[1],
                             np.arange(n[0]) + box_lo[0], indexing="ij")
    return value(c, i, j, k, level)


@pytest.mark.parametrize("dim,n", [(3, (8, 6, 10)), (2, (12, 8, 1))])
def test_cpp_writer_python_reader(driver, tmp_path, dim, n):
    nc = 5
    subprocess.run([driver, "write", str(tmp_path), str(dim), str(nc), *map(str, n)], check=True)
    pf = read_plotfile(str(tmp_path / "plt"))
    assert pf.version == "HyperCLaw-V1.1"
    assert pf.varnames == [f"var{c}" for c in range(nc)]
    assert (pf.dim, pf.time, pf.finest_level, pf.coord, pf.ref_ratio) == (dim, 1.5e-9, 1, 0, [2])
    plo, phi = (-0.01, 0.0, 0.002)[:dim], (0.01, 0.02, 0.034)[:dim]
    assert pf.prob_lo == plo and pf.prob_hi == phi
    for lev in range(2):
        L = pf.levels[lev]
        assert L.domain == ((0,) * dim, tuple(n[d] * 2 ** lev - 1 for d in range(dim)))
        assert L.step == 12 and L.time == 1.5e-9
        dx = tuple((phi[d] - plo[d]) / (n[d] * 2 ** lev) for d in range(dim))
        assert np.allclose(L.dx, dx, rtol=1e-15)
        lo = tuple((n[d] // 2 if lev else 0) for d in range(dim))
        assert L.mf.boxes == [(lo, tuple(lo[d] + n[d] - 1 for d in range(dim)))]
        assert L.mf.ngrow == 0 and L.mf.ncomp == nc
        for d in range(dim):
            assert np.allclose(L.phys_boxes[0][d], (plo[d] + lo[d] * dx[d], plo[d] + (lo[d] + n[d]) * dx[d]),
                               rtol=1e-15, atol=1e-18)
        ex = expected(lo + (0,) * (3 - dim), n, nc, lev)
        assert np.array_equal(L.mf.data, ex)
        assert np.array_equal(L.mf.min[0], ex.reshape(nc, -1).min(axis=1))
        assert np.array_equal(L.mf.max[0], ex.reshape(nc, -1).max(axis=1))
        with open(tmp_path / "plt" / f"Level_{lev}" / "Cell_D_00000", "rb") as f:
            head = f.readline().decode()
        assert YT_FAB.match(head), head
        assert "(8 7 6 5 4 3 2 1)" in head
    assert np.array_equal(pf.field("var3", 1), expected((n[0] // 2, n[1] // 2, n[2] // 2 if dim > 2 else 0),
                                                        n, nc, 1)[3])
    ck = read_checkpoint(str(tmp_path / "chk"), dim)
    assert ck.title == "Checkpoint file for Vidyut"
    assert (ck.finest_level, ck.istep, ck.dt, ck.t_new) == (1, [12, 12], [5e-13, 5e-13], [1.5e-9, 1.5e-9])
    for lev in range(2):
        lo = tuple((n[d] // 2 if lev else 0) for d in range(dim))
        assert ck.boxes[lev] == [(lo, tuple(lo[d] + n[d] - 1 for d in range(dim)))]
        assert np.array_equal(ck.levels[lev].data, expected(lo + (0,) * (3 - dim), n, nc, lev))
    out = subprocess.run([driver, "readck", str(tmp_path / "chk"), str(dim)], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    assert out[0] == "1" and out[1].split() == ["12", "12"]
    assert [float(x) for x in out[2].split()] == [5e-13, 5e-13]
    assert [float(x) for x in out[3].split()] == [1.5e-9, 1.5e-9]


def _write_mf_by_hand(prefix, dim, ncomp, boxes, ngrow, files, big, level):
    """A VisMF MultiFab from the documented layout: header, one FAB per box
    in the named data files (appended in order), ghost cells set to 1e300."""
    base = os.path.dirname(prefix)
    offsets = []
    order = "1 2 3 4 5 6 7 8" if big else "8 7 6 5 4 3 2 1"
    mins, maxs = [], []
    for (lo, hi), fname in zip(boxes, files):
        glo = [lo[d] - ngrow for d in range(dim)]
        ghi = [hi[d] + ngrow for d in range(dim)]
        ext = [ghi[d] - glo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
        vals = np.full((ncomp, ext[2], ext[1], ext[0]), 1e300)
        vlo = list(lo) + [0] * (3 - dim)
        vn = [hi[d] - lo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
        v = expected(vlo, vn, ncomp, level)
        g = [ngrow] * dim + [0] * (3 - dim)
        vals[:, g[2]:g[2] + vn[2], g[1]:g[1] + vn[1], g[0]:g[0] + vn[0]] = v
        mins.append(v.reshape(ncomp, -1).min(axis=1))
        maxs.append(v.reshape(ncomp, -1).max(axis=1))
        path = os.path.join(base, fname)
        off = os.path.getsize(path) if os.path.exists(path) else 0
        offsets.append(off)
        iv = lambda a: "(" + ",".join(str(x) for x in a) + ")"
        head = f"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, ({order})))({iv(glo)} {iv(ghi)} {iv([0] * dim)}) {ncomp}\n"
        with open(path, "ab") as f:
            f.write(head.encode())
            f.write(vals.astype(">f8" if big else "<f8").tobytes())
    iv = lambda a: "(" + ",".join(str(x) for x in a) + ")"
    lines = ["1", "1", str(ncomp), str(ngrow), f"({len(boxes)} 0"]
    lines += [f"({iv(lo)} {iv(hi)} {iv([0] * dim)})" for lo, hi in boxes]
    lines += [")", str(len(boxes))]
    lines += [f"FabOnDisk: {fn} {off}" for fn, off in zip(files, offsets)]
    lines += ["", f"{len(boxes)},{ncomp}"] + ["".join(f"{float(x)!r}," for x in m) for m in mins]
    lines += [f"{len(boxes)},{ncomp}"] + ["".join(f"{float(x)!r}," for x in m) for m in maxs] + [""]
    with open(prefix + "_H", "w") as f:
        f.write("\n".join(lines))


@pytest.mark.parametrize("dim,big,ngrow", [(3, True, 1), (3, False, 2), (2, True, 0)])
def test_decomposed_multifab_read_by_both_readers(driver, tmp_path, dim, big, ngrow):
    nc = 3
    if dim == 3:
        boxes = [((0, 0, 0), (7, 5, 3)), ((0, 0, 4), (7, 5, 9)), ((8, 0, 0), (11, 5, 9))]
        level_box = ((0, 0, 0), (11, 5, 9))
    else:
        boxes = [((0, 0), (5, 7)), ((6, 0), (11, 3)), ((6, 4), (11, 7))]
        level_box = ((0, 0, 0), (11, 7, 0))
    files = ["phi_D_00000", "phi_D_00001", "phi_D_00000"]
    prefix = str(tmp_path / "phi")
    _write_mf_by_hand(prefix, dim, nc, boxes, ngrow, files, big, 0)
    lo, hi = level_box
    n = [hi[d] - lo[d] + 1 for d in range(3)]
    ex = expected(lo, n, nc, 0)
    out = str(tmp_path / "out.bin")
    subprocess.run([driver, "readmf", prefix, str(dim), str(nc), *map(str, lo), *map(str, hi), out], check=True)
    got = np.fromfile(out, dtype="<f8").reshape(nc, n[2], n[1], n[0])
    assert np.array_equal(got, ex)
    mf = read_multifab(prefix, dim)
    assert np.array_equal(mf.data, ex)
    assert mf.ngrow == ngrow and len(mf.boxes) == 3
    for b, (blo, bhi) in enumerate(boxes):
        sub = expected(list(blo) + [0] * (3 - dim), [bhi[d] - blo[d] + 1 for d in range(dim)] + [1] * (3 - dim), nc, 0)
        assert np.array_equal(mf.min[b], sub.reshape(nc, -1).min(axis=1))
        assert np.array_equal(mf.max[b], sub.reshape(nc, -1).max(axis=1))
    # a level box the file does not cover is refused
    bad_hi = (hi[0] + 1,) + tuple(hi[1:])
    r = subprocess.run([driver, "readmf", prefix, str(dim), str(nc), *map(str, lo), *map(str, bad_hi), out],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "do not cover" in r.stderr
