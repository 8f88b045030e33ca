"""Accuracy study of the isentropic vortex (vortex strength 5, centre (5, 5), on the doubly periodic
[0, 10]^2 square): N x N jittered quads (+-10 %) split by random diagonals into 2 N^2 triangles,
Dubiner order p, limiter off, SSP-RK3 at the reference CFL to t_end. For every (N, p) it reports the
L2 error against the exact solution in FP64 and in FP32-bulk, the relative L2 difference between the
two per conserved variable, the observed convergence order between successive N, the step count and the wall time.

python -m dgm_in_rust_amd.studies [--sizes 64,128,256,512] [--orders 1,2,3,4] [--t-end 2.0]
prints one JSON object per run, then a markdown table.

Limiter-order study (--ogrid): the 160 x 50 NACA 0012 O-grid (16 000 triangles), p = 2, M = 0.5,
AoA 1.25, CFL 0.3, marched from freestream for --steps steps with the element-order limiter (the
one-wavefront walk, and one launch per level) and the Jacobi limiter, in FP64 and FP32-bulk. It reports
ms per step (wall clock of the march, upload and download included), GDOF-stage/s, and how far the
walk's final state is from the per-level one.

python -m dgm_in_rust_amd.studies --ogrid [--steps 500] [--modes walk,levels,jacobi] [--precisions fp64,fp32]
                                  [--grid 160x50] [--order 2] [--mach 0.5]

Launch-boundary study (--sync): what separating two passes costs as a boundary between launches of a
captured graph against a grid-wide barrier (cooperative_groups grid.sync()) inside one cooperative
launch, for light passes (one float per thread) of 256 threads per block and several grid sizes. For
each size it times `phases` passes and one pass both ways and reports the cost of one separation,
(T(phases) - T(1)) / (phases - 1), from the kernels of libdgm_study.so.

python -m dgm_in_rust_amd.studies --sync [--blocks 256,512,1024,2048] [--phases 9] [--reps 200]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import time

import numpy as np

from . import cases, meshgen
from .solver import Context, SpatialDisc, TemperalDisc


def vortex_run(n: int, order: int, precision: str, t_end: float, seed: int = 1):
    mesh = meshgen.periodic_square(n, jitter=0.1, seed=seed)
    ctx = Context(mesh, order=order, limiter="off", mach=0.5, aoa_deg=0.0, precision=precision)
    U = cases.vortex_state(ctx, 0.0)
    t0 = time.perf_counter()
    st = TemperalDisc(ctx, final_step=10**8, final_time=t_end).time_march(SpatialDisc(ctx), ctx.zeros(), U)
    wall = time.perf_counter() - t0
    err = cases.vortex_error(ctx, U, t_end)
    return ctx, U, dict(n=n, order=order, precision=precision, triangles=2 * n * n, steps=int(st["steps"]),
                        time=float(st["time"]), wall_s=wall, l2_density=float(err[0]),
                        l2_all=float(np.sqrt((err * err).sum())))


# limiter-order study modes: (limiter, value of DGM_SEQUENTIAL_WALK read at context creation)
ORDER_MODES = {"walk": ("sequential", "1"), "levels": ("sequential", "0"), "jacobi": ("jacobi", "1")}


def ogrid_order_study(steps: int, modes, precisions, surface=160, layers=50, order=2, mach=0.5):
    mesh = meshgen.naca0012_ogrid(surface, layers, seed=0)
    rows = []
    saved = os.environ.get("DGM_SEQUENTIAL_WALK")
    try:
        _order_rows(mesh, steps, modes, precisions, order, mach, rows)
    finally:
        # the walk choice is read at context creation: later contexts of this process get the caller's
        if saved is None:
            os.environ.pop("DGM_SEQUENTIAL_WALK", None)
        else:
            os.environ["DGM_SEQUENTIAL_WALK"] = saved
    return rows


def _order_rows(mesh, steps, modes, precisions, order, mach, rows):
    for prec in precisions:
        final = {}
        for mode in modes:
            lim, walk = ORDER_MODES[mode]
            os.environ["DGM_SEQUENTIAL_WALK"] = walk
            ctx = Context(mesh, order=order, cfl=0.3, mach=mach, aoa_deg=1.25, precision=prec, limiter=lim)
            U0 = ctx.initial_solution(1.0, 1.25, mach, 1.0e5)
            U = U0.copy()  # warm-up: graph capture and first launches
            TemperalDisc(ctx, final_step=20, final_time=1e9).time_march(SpatialDisc(ctx), ctx.zeros(), U)
            U = U0.copy()
            t0 = time.perf_counter()
            st = TemperalDisc(ctx, final_step=steps, final_time=1e9).time_march(SpatialDisc(ctx), ctx.zeros(), U)
            wall = time.perf_counter() - t0
            n = int(st["steps"])
            row = dict(precision=prec, mode=mode, steps=n, ms_per_step=wall / n * 1e3,
                       gdof_stage_per_s=ctx.ne * 4 * ctx.nbasis * 3 * n / wall / 1e9)
            final[mode] = U
            rows.append(row)
            print(json.dumps(row), flush=True)
            ctx.close()
        os.environ.pop("DGM_SEQUENTIAL_WALK", None)
        if "walk" in final and "levels" in final:
            d = float(np.abs(final["walk"] - final["levels"]).max())
            print(json.dumps(dict(precision=prec, walk_vs_levels_max_abs=d)), flush=True)
    return rows


def sync_study(blocks_list, phases: int, reps: int, threads: int = 256):
    """Cost of one separation between passes: a launch boundary inside a graph and a grid barrier inside a
    cooperative launch (libdgm_study.so). Grid sizes that cannot be co-resident are reported and skipped."""
    import ctypes as C
    lib = C.CDLL(os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdgm_study.so"))
    lib.dgm_study_sync.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    lib.dgm_study_max_grid.argtypes = [C.c_int]
    maxg = lib.dgm_study_max_grid(threads)
    rows = []
    for nb in blocks_list:
        if nb > maxg:
            print(json.dumps(dict(blocks=nb, skipped=f"more than the {maxg} co-resident blocks")), flush=True)
            continue
        t = {}
        for ph in (1, phases):
            g, c = C.c_double(), C.c_double()
            st = lib.dgm_study_sync(nb, threads, ph, reps, C.byref(g), C.byref(c))
            if st:
                raise RuntimeError(f"dgm_study_sync({nb}, {threads}, {ph}, {reps}) returned {st}")
            t[ph] = (g.value * 1e3, c.value * 1e3)   # microseconds
        row = dict(blocks=nb, threads=threads, phases=phases, reps=reps, max_coresident_blocks=maxg,
                   graph_us_1=t[1][0], graph_us_n=t[phases][0], grid_us_1=t[1][1], grid_us_n=t[phases][1],
                   launch_boundary_us=(t[phases][0] - t[1][0]) / (phases - 1),
                   grid_barrier_us=(t[phases][1] - t[1][1]) / (phases - 1))
        rows.append(row)
        print(json.dumps(row), flush=True)
    print(f"\n| Blocks (x {threads} threads) | {phases} passes in a graph (us) | one launch boundary (us) | "
          f"{phases} passes in one cooperative launch (us) | one grid barrier (us) |")
    print("|---|---|---|---|---|")
    for r in rows:
        print(f"| {r['blocks']} | {r['graph_us_n']:.1f} | {r['launch_boundary_us']:.1f} | {r['grid_us_n']:.1f} | "
              f"{r['grid_barrier_us']:.1f} |")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="64,128,256,512")
    ap.add_argument("--orders", default="1,2,3,4")
    ap.add_argument("--t-end", type=float, default=2.0)
    ap.add_argument("--ogrid", action="store_true", help="the limiter-order study instead of the vortex")
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--modes", default="walk,levels,jacobi")
    ap.add_argument("--precisions", default="fp64,fp32")
    ap.add_argument("--grid", default="160x50", help="O-grid surface points x layers (--ogrid)")
    ap.add_argument("--order", type=int, default=2, help="polynomial order (--ogrid)")
    ap.add_argument("--mach", type=float, default=0.5, help="freestream Mach number (--ogrid)")
    ap.add_argument("--sync", action="store_true", help="the launch-boundary study instead of the vortex")
    ap.add_argument("--blocks", default="256,512,1024,2048", help="grid sizes (--sync)")
    ap.add_argument("--phases", type=int, default=9, help="passes per measurement (--sync)")
    ap.add_argument("--reps", type=int, default=200, help="repetitions per measurement (--sync)")
    a = ap.parse_args(argv)
    if a.sync:
        sync_study([int(x) for x in a.blocks.split(",")], a.phases, a.reps)
        return
    if a.ogrid:
        ns, nl = (int(x) for x in a.grid.split("x"))
        ogrid_order_study(a.steps, a.modes.split(","), a.precisions.split(","), ns, nl, a.order, a.mach)
        return
    sizes = [int(x) for x in a.sizes.split(",")]
    orders = [int(x) for x in a.orders.split(",")]
    rows = []
    for p in orders:
        prev = None
        for n in sizes:
            c64, U64, r64 = vortex_run(n, p, "fp64", a.t_end)
            c32, U32, r32 = vortex_run(n, p, "fp32", a.t_end)
            row = dict(r64)
            pv = cases.relative_l2_per_variable(c64, U32, U64)
            row.update(l2_density_fp32=r32["l2_density"], l2_all_fp32=r32["l2_all"], wall_fp32_s=r32["wall_s"],
                       steps_fp32=r32["steps"], fp32_vs_fp64_rel_l2=cases.relative_l2(c64, U32, U64),
                       fp32_vs_fp64_per_variable=pv, fp32_vs_fp64_max_variable=max(pv.values()))
            row["rate"] = None if prev is None else math.log2(prev / row["l2_density"])
            prev = row["l2_density"]
            rows.append(row)
            print(json.dumps(row), flush=True)
    print("\n| N | p | triangles | steps | L2 err (rho) FP64 | L2 err (rho) FP32-bulk | rate | FP32 vs FP64 rel. L2: rho, rho u, rho v, E | wall FP64 / FP32 (s) |")
    print("|---|---|---|---|---|---|---|---|---|")
    for r in rows:
        rate = "" if r["rate"] is None else f"{r['rate']:.2f}"
        pv = ", ".join(f"{v:.1e}" for v in r["fp32_vs_fp64_per_variable"].values())
        print(f"| {r['n']} | {r['order']} | {r['triangles']} | {r['steps']} | {r['l2_density']:.3e} | "
              f"{r['l2_density_fp32']:.3e} | {rate} | {pv} | {r['wall_s']:.1f} / {r['wall_fp32_s']:.1f} |")


if __name__ == "__main__":
    main()
