"""Benchmark of the vidyut3d per-step plasma-fluid path on one MI355X.

Workload: the 3D LFA air streamer, 256^3 cells over a 2 cm cube, 10 species
(8 transported), 20 E/N-Arrhenius reactions, WENO5-Z fluxes, multigrid
Poisson to 1e-10 relative residual, dt = 5e-13 s (cases/air3d).  A step is
one full Evolve step (time scales, Poisson solve + E field, fluxes +
reactions, batched implicit species solves) on device-resident state.

Prints ONE JSON line: Mcell-steps/s (FP64), the dominant kernel's roofline
from HIP-event timings taken inside the timed region, and the oracle (CPU
restatement) timed on a bounded sample on this host as cpu_baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s: FP64 plain vs mixed @1e-8 rel err; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def kernel_bytes_per_cell(name: str, S: int, nt: int, D: int, c: int = 8) -> float | None:
    """Algorithmic (compulsory) HBM bytes per finest-level cell per launch:
    each array the kernel reads, read once, and each it writes, written once,
    for the air3d case, whose mobilities and diffusivities are constants (its
    specMob/specDiff read nothing from the state, so the kernels read no
    density sum, |E| or T_e for them).  S species (with the two background
    gases), nt transported, D dimensions, c bytes per value."""
    q = 1.0 / (1 << D)
    Sq = S - 3          # charged species: all but N2*, N2, O2
    table = {
        # one direction: nt densities and E_d read, nt faces written
        "flux": (2 * nt + 1) * c,
        # S densities (rates), T_e and E/N, D*nt faces read; nt rhs written
        # (the old densities are the same array at one corrector; the 1/D of
        # air3d's uniform diffusivities is not stored after the first step)
        "cell_source": (S + 2 + D * nt + nt) * c,
        "poisson_rhs": (Sq + 1) * c,
        # solution and S densities (E/N) read; POT, E (D) and E/N written
        "efield": (1 + S + 1 + D + 1) * c,
        # one colour of a red-black sweep: x read (neighbours), half f read, half x written
        "pois_gsrb_l0": 2 * c,
        "pois_resid_l0": 3 * c,
        "pois_resid_restrict_l0": (2 + q) * c,
        "pois_prolong_l0": (2 + q) * c,
        # E (D components) and the electron density
        "time_scales": (D + 1) * c,
    }
    return table.get(name)


def kernel_flops_per_cell(name: str, S: int, nt: int) -> float | None:
    """Algorithmic FP64 flops per finest-level cell (face) per launch for the
    compute-bound WENO5 flux: per face, 6 stencil-cell reductions (sum of S
    densities, |E|: ~15 flops each), and per transported species 6 drift
    velocities and products, four WENO-Z reconstructions (~75 flops each,
    a division counted as one flop) and the Lax-Friedrichs combination."""
    if name == "flux":
        return 6 * 15 + nt * (12 + 4 * 75 + 8)
    return None


def lmea_leg(steps: int, warmup: int, n: int, levels: int) -> dict:
    """The 3D LMEA argon case with the electron energy equation (lmea3d):
    base n^3 plus `levels` nested levels of ratio 2, each the central 1/8 of
    the level below, RF 300 V at 13.56 MHz, dt = 1e-11 s; every level
    advances every step.  Throughput counts the cells of all levels."""
    from vidyut3d_amd import Vidyut, read_inputs

    v = Vidyut("lmea3d", inputs=read_inputs("lmea3d", {"amr.n_cell": f"{n} {n} {n}", "amr.max_level": levels}))
    v.init_state()
    if warmup:
        v.evolve(warmup)
    v.synchronize()
    t0 = time.perf_counter()
    info = v.evolve(steps)
    wall = time.perf_counter() - t0
    cells = 0
    for lev in range(v.num_levels()):
        _, ext = v.level_box(lev)
        cells += ext[0] * ext[1] * ext[2]
    v.close()
    return {
        "workload": f"lmea3d {n}^3 + {levels} levels (ratio 2, central 1/8 each), 5 species + electron energy, "
                    f"RF 300 V at 13.56 MHz, dt 1e-11, FP64",
        "cells_all_levels": cells, "steps": steps, "warmup": warmup,
        "value": round(cells * steps / wall / 1e6, 3), "unit": "Mcell-steps/s",
        "ms_per_step": round(wall * 1e3 / steps, 3),
        "poisson_vcycles_last_step": info.poisson_iters, "species_vcycles_last_step": info.species_iters,
    }


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "air3d_256_hbm_traffic.md")


def hbm_traffic(kernel: str) -> dict | None:
    """HBM bytes per launch of a kernel class as measured with rocprofv3
    FETCH_SIZE / WRITE_SIZE counters (scaled by a plain copy of known size)
    and committed in profiles/air3d_256_hbm_traffic.md: the mean over the
    class's launches of one step and, for the flux, per direction; None if
    absent."""
    if kernel != "flux" or not os.path.exists(TRAFFIC_PROFILE):
        return None
    out = {}
    for line in open(TRAFFIC_PROFILE):
        for key, name in (("flux HBM bytes per launch, x", "x"), ("flux HBM bytes per launch, y", "y"),
                          ("flux HBM bytes per launch, z", "z")):
            if line.startswith(key):
                out[name] = float(line.rsplit(":", 1)[1])
    if not out:
        return None
    out["mean"] = sum(out.values()) / len(out)
    return out


def _host_cpu() -> tuple[str, set]:
    model, flags = "unknown", set()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name") and model == "unknown":
                model = line.split(":", 1)[1].strip()
            if line.startswith("flags") and not flags:
                flags = set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return model, flags


def cpu_baseline(steps: int = 1, n: int = 256) -> dict:
    """The oracle (CPU restatement, OpenMP over all host threads) on a
    bounded sample of the same workload: air3d at n^3 for `steps` timed steps
    from the initial state, setup excluded (a first step measured no slower
    than a second, so no warm-up step is spent).  It runs the oracle built -O3
    for x86-64-v3 (AVX2/FMA) when the host has those instructions, else the
    -O2 build the parity tests use.  Reported as a baseline, not a target."""
    from oracle import Oracle, load
    from oracle.build import FAST_FLAGS
    from vidyut3d_amd import read_inputs

    model, flags = _host_cpu()
    fast = {"avx2", "fma", "bmi2", "movbe", "f16c", "abm"} <= flags
    variant = "fast" if fast else ""
    lib = load("air3d", variant=variant)
    o = Oracle("air3d", read_inputs("air3d", {"amr.n_cell": f"{n} {n} {n}"}), variant=variant)
    t0 = time.perf_counter()
    o.evolve(steps)
    dt = time.perf_counter() - t0
    o.close()
    return {
        "value": round(n ** 3 * steps / dt / 1e6, 4),
        "unit": "Mcell-steps/s",
        "cores": int(lib.or_num_threads()),
        "kind": "port",
        "sample": f"air3d {n}^3 (the benchmark grid), {steps} step(s) from the initial state ({dt:.1f} s), "
                  f"oracle C++/OpenMP",
        "cpu_model": model,
        "compiler_flags": FAST_FLAGS if fast else "g++ -O2 -fopenmp",
    }


def adjoint_leg(steps: int, every: int, n: int) -> dict:
    """The RF argon case (argonrf2d): forward objective (time-integrated
    electrode current, checkpoints every `every` steps) against the
    hand-written adjoint (segment recompute + reverse steps), both timed on
    the host clock around calls that synchronise."""
    import numpy as np
    from vidyut3d_amd import Vidyut, read_inputs

    v = Vidyut("argonrf2d", inputs=read_inputs("argonrf2d", {"amr.n_cell": f"{n} {n}"}))
    v.init_state()
    v.objective_forward(every, every)      # warm-up (code objects, buffers)
    v.adjoint_gradient()
    v.init_state()
    v.synchronize()
    t0 = time.perf_counter()
    J = v.objective_forward(steps, every)
    tf = time.perf_counter() - t0
    t0 = time.perf_counter()
    g = v.adjoint_gradient()
    ta = time.perf_counter() - t0
    p = v.get_params()
    v.close()
    return {
        "workload": f"argonrf2d {n}^2, {steps} steps x 1e-11 s, RF 500 V at 13.56 MHz, "
                    f"objective = time-integrated x-hi electrode current, checkpoints every {every}",
        "forward_s": round(tf, 4), "adjoint_s": round(ta, 4), "adjoint_over_forward": round(ta / tf, 3),
        "forward_Mcell_steps_per_s": round(n * n * steps / tf / 1e6, 3),
        "J": J, "n_params": int(len(g)), "max_abs_dJdlogp": float(np.max(np.abs(g * p))),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per direction")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint leg")
    ap.add_argument("--no-direct", action="store_true",
                    help="skip the leg with the finest-grid direct Poisson solve")
    ap.add_argument("--adj-steps", type=int, default=400)
    ap.add_argument("--adj-every", type=int, default=20)
    ap.add_argument("--adj-n", type=int, default=512)
    ap.add_argument("--no-lmea", action="store_true", help="skip the multi-level LMEA leg")
    ap.add_argument("--lmea-n", type=int, default=192)
    ap.add_argument("--lmea-levels", type=int, default=2)
    ap.add_argument("--lmea-steps", type=int, default=10)
    ap.add_argument("--mixed-n2", type=int, default=384,
                    help="second grid for the FP64 vs mixed comparison (0: skip)")
    ap.add_argument("--mixed-n2-steps", type=int, default=10)
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="inputs override for the air3d legs (ParmParse key, repeatable)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB 4 times (calibrates HBM byte counters)")
    args = ap.parse_args()

    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(4):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 8, "copy_probe_bytes_written": n * 8, "launches": 4}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (--gpus {args.gpus} requested); "
              "the multi-GPU decomposition is a design, not a runnable path", file=sys.stderr)
        sys.exit(2)

    from vidyut3d_amd import Vidyut, read_inputs

    n = args.n

    def leg(mixed: bool, n: int = n, steps: int = args.steps, time_dominant: bool = False,
            extra: dict | None = None):
        ov = {"amr.n_cell": f"{n} {n} {n}", "vidyut.mixed_precision": int(mixed)}
        ov.update(extra or {})
        ov.update(dict(kv.split("=", 1) for kv in args.set))
        v = Vidyut("air3d", inputs=read_inputs("air3d", ov))
        v.init_state()
        if args.warmup:
            v.evolve(args.warmup)
        # per-class breakdown from an instrumented (untimed) pass of 2 steps
        v.kernel_timing(True)
        v.evolve(2)
        kt = v.kernel_times()
        # the dominant kernel: the class with the most time among those of a
        # few launches per step (one kernel per launch, not a solver's many
        # coarse-level launches)
        dom = max(((k, t) for k, t in kt.items() if 0 < t[1] <= 32), key=lambda kv: kv[1][0])[0]
        # timed region from the initial state: HIP events only around the
        # dominant kernel's launches, and only where its roofline is taken
        v.init_state()
        v.set_time(0.0)
        v.kernel_timing(time_dominant, [dom] if time_dominant else None)
        v.synchronize()
        t0 = time.perf_counter()
        info = v.evolve(steps)   # synchronises at the end
        wall = time.perf_counter() - t0
        kt_dom = v.kernel_times()[dom] if time_dominant else None
        v.kernel_timing(False)
        return v, wall, info, kt, dom, kt_dom

    v, wall, info, kt, dom, kt_dom = leg(False, time_dominant=True)
    d = v.dims
    ncells = d.ncells
    ms_per_step = wall * 1e3 / args.steps
    value = ncells * args.steps / wall / 1e6

    S, nt, D = d.nspec, d.ntransported, d.dim
    kernels = {}
    for name, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        e = {"ms_per_step": round(ms / 2, 4), "launches_per_step": round(cnt / 2, 2),
             "avg_us": round(ms * 1e3 / cnt, 2)}
        b = kernel_bytes_per_cell(name, S, nt, D)
        if b is not None:
            e["GBps"] = round(b * ncells / (ms / cnt * 1e-3) / 1e9, 1)
        kernels[name] = e
    dms, dcnt = kt_dom
    avg_s = dms / dcnt * 1e-3
    b = kernel_bytes_per_cell(dom, S, nt, D)
    fl = kernel_flops_per_cell(dom, S, nt)
    if fl is not None:
        ach = fl * ncells / avg_s / 1e12
        # FP64 vector (VALU) work, no matrix cores: priced against MI355X's
        # FP64 vector peak
        tr = hbm_traffic(dom)
        roof = {"kernel": dom, "bound": "fp64_valu", "achieved": round(ach, 2), "peak": FP64_PEAK_TFS,
                "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFS, 4),
                "traffic": tr["mean"] if tr else None,
                "traffic_per_direction": {k: tr[k] for k in ("x", "y", "z") if k in tr} if tr else None,
                "algorithmic_bytes_per_launch": b * ncells if b else None,
                "flops_per_launch": fl * ncells, "avg_launch_us": round(avg_s * 1e6, 2),
                "hbm_GBps": round(b * ncells / avg_s / 1e9, 1) if b else None}
    elif b is not None:
        ach = b * ncells / avg_s / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": b * ncells, "avg_launch_us": round(avg_s * 1e6, 2)}
    else:
        roof = {"kernel": dom, "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": None, "traffic": None, "avg_launch_us": round(avg_s * 1e6, 2)}

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"air3d streamer {n}^3, 10 species (8 transported), 20 reactions, "
                               f"WENO5-Z, MG Poisson reltol 1e-10, dt 5e-13, FP64 plain",
                   "cells": ncells, "poisson_vcycles_last_step": info.poisson_iters,
                   "species_vcycles_last_step": info.species_iters},
        "roofline": roof,
        "kernels": kernels,
    }
    def field_errors(a, b):
        """Per-field max and L2 relative differences of state a against b
        (transported species, potential, E/N, E)."""
        import numpy as np
        fields = {}
        worst_max = worst_l2 = 0.0
        names = {d.nspec + 2: "POT", d.nspec + 9: "REF"}
        for c in list(range(d.ntransported)) + [d.nspec + 2, d.nspec + 9]:
            den = float(np.max(np.abs(b[c])))
            mx = float(np.max(np.abs(a[c] - b[c]))) / den if den > 0 else 0.0
            l2 = float(np.linalg.norm(a[c] - b[c]) / max(np.linalg.norm(b[c]), 1e-300))
            fields[names.get(c, f"species{c}")] = {"max_rel": mx, "l2_rel": l2}
            worst_max, worst_l2 = max(worst_max, mx), max(worst_l2, l2)
        efx = d.nspec + 3
        mag = np.sqrt(sum(b[efx + q] ** 2 for q in range(d.dim)))
        e_max = max(float(np.max(np.abs(a[efx + q] - b[efx + q]))) for q in range(d.dim)) / float(np.max(mag))
        e_l2 = float(np.sqrt(sum(np.sum((a[efx + q] - b[efx + q]) ** 2) for q in range(d.dim))) /
                     np.sqrt(np.sum(mag ** 2)))
        fields["E"] = {"max_rel": e_max, "l2_rel": e_l2}
        worst_max, worst_l2 = max(worst_max, e_max), max(worst_l2, e_l2)
        return fields, worst_max, worst_l2

    if not args.no_direct:
        # the same FP64 steps with the Poisson equation of each step solved
        # directly on the finest grid (separable operator: the eigen-basis
        # products as GEMMs), then the same stopping test; fields against the
        # V-cycle leg's
        vd_, walld, infod, ktd, _, _ = leg(False, extra={"vidyut.mg_direct_finest": 1,
                                                        "vidyut.mg_direct_bottom": n})
        fields_d, wmax_d, wl2_d = field_errors(vd_.get_state(), v.get_state())
        out["poisson_direct"] = {
            "workload": f"air3d {n}^3, {args.steps} steps, FP64, vidyut.mg_direct_finest = 1 "
                        f"(Poisson by one direct solve of the {n}^3 grid + the reltol 1e-10 test)",
            "value": round(ncells * args.steps / walld / 1e6, 3), "unit": "Mcell-steps/s",
            "ms_per_step": round(walld * 1e3 / args.steps, 3),
            "speedup_vs_vcycles": round(wall / walld, 3),
            "poisson_iters_last_step": infod.poisson_iters,
            "field_err_vs_vcycles": {"worst_max_rel": wmax_d, "worst_l2_rel": wl2_d, "per_field": fields_d},
            "kernels": {k: {"ms_per_step": round(ms / 2, 4), "launches_per_step": round(cnt / 2, 2)}
                        for k, (ms, cnt) in ktd.items() if cnt},
        }
        vd_.close()
    if not args.no_mixed:
        v32, wall32, info32, kt32, dom32, _ = leg(True)
        fields, worst_max, worst_l2 = field_errors(v32.get_state(), v.get_state())
        out["mixed"] = {
            "value": round(ncells * args.steps / wall32 / 1e6, 3), "unit": "Mcell-steps/s",
            "ms_per_step": round(wall32 * 1e3 / args.steps, 3),
            "speedup_vs_fp64": round(wall / wall32, 3),
            "poisson_vcycles_last_step": info32.poisson_iters,
            "field_err_vs_fp64": {"worst_max_rel": worst_max, "worst_l2_rel": worst_l2, "per_field": fields},
            "kernels": {k: {"ms_per_step": round(ms / 2, 4), "launches_per_step": round(cnt / 2, 2)}
                        for k, (ms, cnt) in kt32.items() if cnt},
            "dtype": "f32 Poisson V-cycles in f64 refinement + f32 WENO5 faces; f64 elsewhere",
        }
        v32.close()
    if not args.no_mixed and args.mixed_n2:
        # the same comparison on a larger grid (bandwidth scaling)
        n2, k2 = args.mixed_n2, args.mixed_n2_steps
        w2 = {}
        for mixed in (False, True):
            v2, wall2, info2, _, _, _ = leg(mixed, n2, k2)
            w2[mixed] = wall2
            v2.close()
        out["mixed_n2"] = {
            "workload": f"air3d {n2}^3, {k2} steps", "fp64_value": round(n2 ** 3 * k2 / w2[False] / 1e6, 3),
            "mixed_value": round(n2 ** 3 * k2 / w2[True] / 1e6, 3), "unit": "Mcell-steps/s",
            "speedup_vs_fp64": round(w2[False] / w2[True], 3),
        }
    if not args.no_lmea:
        out["lmea_levels"] = lmea_leg(args.lmea_steps, 2, args.lmea_n, args.lmea_levels)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args.adj_steps, args.adj_every, args.adj_n)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_steps, args.cpu_n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
