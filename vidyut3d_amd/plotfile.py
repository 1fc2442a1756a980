"""Readers for AMReX plotfiles and the reference's checkpoints
(Source/Plot.cpp:32-225), independent of the library's C++ writer.

A plotfile directory holds a text Header (format name "HyperCLaw-V1.1",
variable names, dimension, time, finest level, problem extents, refinement
ratios, level domains, level steps, cell sizes, coordinate system, then per
level its boxes in physical coordinates and the path of its MultiFab) and
one VisMF MultiFab per level (Level_<l>/Cell_H plus data files).  A
checkpoint holds the reference's Header (title, finest level, istep, dt and
t_new per level, one BoxArray per level) and Level_<l>/phi.  A VisMF header
lists, per box, the data file and the byte offset of a line
"FAB <real descriptor><box> <ncomp>" followed by the box's values (ghost
cells included), x fastest, component outermost; the descriptor's second
part gives the byte order ((8 7 6 5 4 3 2 1) little-endian,
(1 2 3 4 5 6 7 8) big-endian).

These read decomposed domains (many boxes, many data files), so the
reference's own output can be read the same way.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass, field

import numpy as np

_INT = re.compile(r"-?\d+")


def _ints(s: str) -> list[int]:
    return [int(x) for x in _INT.findall(s)]


def _box(s: str, dim: int) -> tuple[tuple[int, ...], tuple[int, ...]]:
    v = _ints(s)
    if len(v) < 2 * dim:
        raise ValueError(f"malformed box {s!r}")
    return tuple(v[:dim]), tuple(v[dim:2 * dim])


def _read_boxarray(lines: list[str], i: int, dim: int):
    while "(" not in lines[i]:
        i += 1
    n = _ints(lines[i])[0]
    boxes = [_box(lines[i + 1 + k], dim) for k in range(n)]
    return boxes, i + n + 2          # past the closing ")"


@dataclass
class MultiFab:
    """A MultiFab read onto the bounding box of its valid boxes:
    data[c, k, j, i] (2D: k has extent 1), NaN where no box covers."""
    ncomp: int
    ngrow: int
    boxes: list
    lo: tuple
    hi: tuple
    data: np.ndarray
    min: np.ndarray = None   # per box, per component (header)
    max: np.ndarray = None


_FAB = re.compile(r"^FAB \(\((\d+), \(([0-9 ]+)\)\),\((\d+), \(([0-9 ]+)\)\)\)(\(.*\)) (\d+)\s*$")


def read_multifab(prefix: str, dim: int) -> MultiFab:
    """VisMF::Read: <prefix>_H and the data files it names."""
    with open(prefix + "_H") as f:
        lines = f.read().split("\n")
    vers, how, ncomp = int(lines[0]), int(lines[1]), int(lines[2])
    if vers != 1:
        raise ValueError(f"{prefix}_H: VisMF header version {vers} is not read")
    ng = _ints(lines[3])
    ngrow = ng[0] if ng else 0
    boxes, i = _read_boxarray(lines, 4, dim)
    while not lines[i].strip():
        i += 1
    nfod = int(lines[i])
    fod = []
    for k in range(nfod):
        tag, name, off = lines[i + 1 + k].split()
        if tag != "FabOnDisk:":
            raise ValueError(f"{prefix}_H: malformed FabOnDisk line")
        fod.append((name, int(off)))
    i += 1 + nfod
    # per-box minima then maxima: a line "nboxes,ncomp", then one line of
    # comma-terminated values per box
    rest = [ln for ln in lines[i:] if ln.strip()]

    def table(at):
        nb, nc = (int(x) for x in rest[at].split(","))
        rows = [[float(x) for x in rest[at + 1 + b].split(",") if x.strip()] for b in range(nb)]
        return np.array(rows).reshape(nb, nc), at + 1 + nb

    mins = maxs = None
    if rest:
        mins, at = table(0)
        maxs, _ = table(at)
    lo = tuple(min(b[0][d] for b in boxes) for d in range(dim))
    hi = tuple(max(b[1][d] for b in boxes) for d in range(dim))
    ext = [hi[d] - lo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
    data = np.full((ncomp, ext[2], ext[1], ext[0]), np.nan)
    base = os.path.dirname(prefix)
    for (blo, bhi), (name, off) in zip(boxes, fod):
        with open(os.path.join(base, name), "rb") as f:
            f.seek(off)
            head = f.readline().decode()
            m = _FAB.match(head)
            if not m:
                raise ValueError(f"{name}: malformed FAB header {head!r}")
            nbytes, order = int(m.group(3)), m.group(4).split()
            if int(m.group(1)) != 8 or nbytes != 8:
                raise ValueError(f"{name}: only 8-byte reals are read")
            if order == [str(k) for k in range(8, 0, -1)]:
                dt = np.dtype("<f8")
            elif order == [str(k) for k in range(1, 9)]:
                dt = np.dtype(">f8")
            else:
                raise ValueError(f"{name}: byte order {order}")
            flo, fhi = _box(m.group(5), dim)
            if int(m.group(6)) != ncomp:
                raise ValueError(f"{name}: FAB component count")
            fe = [fhi[d] - flo[d] + 1 for d in range(dim)] + [1] * (3 - dim)
            n = fe[0] * fe[1] * fe[2]
            vals = np.frombuffer(f.read(8 * n * ncomp), dtype=dt).astype(np.float64)
        if vals.size != n * ncomp:
            raise ValueError(f"{name}: truncated data")
        fab = vals.reshape(ncomp, fe[2], fe[1], fe[0])
        g = [ngrow] * dim + [0] * (3 - dim)
        sl_src = tuple(slice(g[d], g[d] + (bhi[d] - blo[d] + 1)) for d in (2, 1, 0) if d < dim)
        sl_dst = tuple(slice(blo[d] - lo[d], bhi[d] - lo[d] + 1) for d in (2, 1, 0) if d < dim)
        if dim == 2:
            sl_src = (slice(0, 1),) + sl_src
            sl_dst = (slice(0, 1),) + sl_dst
        data[(slice(None),) + sl_dst] = fab[(slice(None),) + sl_src]
    return MultiFab(ncomp, ngrow, boxes, lo, hi, data, mins, maxs)


@dataclass
class PlotLevel:
    domain: tuple
    dx: tuple
    step: int
    time: float
    phys_boxes: list
    mf: MultiFab


@dataclass
class PlotFile:
    version: str
    varnames: list
    dim: int
    time: float
    finest_level: int
    prob_lo: tuple
    prob_hi: tuple
    ref_ratio: list
    coord: int
    level_steps: list
    levels: list = field(default_factory=list)

    def field(self, name: str, level: int = 0) -> np.ndarray:
        """One variable of a level on the bounding box of its boxes."""
        return self.levels[level].mf.data[self.varnames.index(name)]


def read_plotfile(path: str) -> PlotFile:
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    version = lines[0].strip()
    nvar = int(lines[1])
    names = [lines[2 + k].strip() for k in range(nvar)]
    i = 2 + nvar
    dim = int(lines[i])
    time = float(lines[i + 1])
    finest = int(lines[i + 2])
    plo = tuple(float(x) for x in lines[i + 3].split())
    phi = tuple(float(x) for x in lines[i + 4].split())
    rr = [int(x) for x in lines[i + 5].split()]
    doms = re.findall(r"\(\([^()]*\) \([^()]*\) \([^()]*\)\)", lines[i + 6])
    domains = [_box(d, dim) for d in doms]
    steps = [int(x) for x in lines[i + 7].split()]
    dxs = [tuple(float(x) for x in lines[i + 8 + l].split()) for l in range(finest + 1)]
    i += 8 + finest + 1
    coord = int(lines[i])
    i += 2                                # coordinate system, boundary width
    pf = PlotFile(version, names, dim, time, finest, plo, phi, rr, coord, steps)
    for lev in range(finest + 1):
        lv, ngrids, t = lines[i].split()
        step = int(lines[i + 1])
        i += 2
        pboxes = []
        for _ in range(int(ngrids)):
            pboxes.append(tuple(tuple(float(x) for x in lines[i + d].split()) for d in range(dim)))
            i += dim
        mfpath = lines[i].strip()
        i += 1
        mf = read_multifab(os.path.join(path, mfpath), dim)
        pf.levels.append(PlotLevel(domains[lev], dxs[lev], step, float(t), pboxes, mf))
    return pf


@dataclass
class Checkpoint:
    title: str
    finest_level: int
    istep: list
    dt: list
    t_new: list
    boxes: list                           # per level
    levels: list = field(default_factory=list)   # MultiFab per level


def read_checkpoint(path: str, dim: int) -> Checkpoint:
    """The reference's checkpoint: its Header and Level_<l>/phi."""
    with open(os.path.join(path, "Header")) as f:
        lines = f.read().split("\n")
    finest = int(lines[1])
    istep = [int(x) for x in lines[2].split()]
    dt = [float(x) for x in lines[3].split()]
    t_new = [float(x) for x in lines[4].split()]
    i = 5
    boxes = []
    for _ in range(finest + 1):
        ba, i = _read_boxarray(lines, i, dim)
        boxes.append(ba)
    ck = Checkpoint(lines[0], finest, istep, dt, t_new, boxes)
    for lev in range(finest + 1):
        ck.levels.append(read_multifab(os.path.join(path, f"Level_{lev}", "phi"), dim))
    return ck
