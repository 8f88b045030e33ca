"""Accuracy study of the isentropic vortex (vortex strength 5, centre (5, 5), on the doubly periodic
[0, 10]^2 square): N x N jittered quads (+-10 %) split by random diagonals into 2 N^2 triangles,
Dubiner order p, limiter off, SSP-RK3 at the reference CFL to t_end. For every (N, p) it reports the
L2 error against the exact solution in FP64 and in FP32-bulk, the relative L2 difference between the
two, the observed convergence order between successive N, the step count and the wall time.

python -m dgm_in_rust_amd.studies [--sizes 64,128,256,512] [--orders 1,2,3,4] [--t-end 2.0]
prints one JSON object per run, then a markdown table.
"""
from __future__ import annotations

import argparse
import json
import math
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


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="64,128,256,512")
    ap.add_argument("--orders", default="1,2,3,4")
    ap.add_argument("--t-end", type=float, default=2.0)
    a = ap.parse_args(argv)
    sizes = [int(x) for x in a.sizes.split(",")]
    orders = [int(x) for x in a.orders.split(",")]
    rows = []
    for p in orders:
        prev = None
        for n in sizes:
            c64, U64, r64 = vortex_run(n, p, "fp64", a.t_end)
            c32, U32, r32 = vortex_run(n, p, "fp32", a.t_end)
            row = dict(r64)
            row.update(l2_density_fp32=r32["l2_density"], l2_all_fp32=r32["l2_all"], wall_fp32_s=r32["wall_s"],
                       steps_fp32=r32["steps"], fp32_vs_fp64_rel_l2=cases.relative_l2(c64, U32, U64))
            row["rate"] = None if prev is None else math.log2(prev / row["l2_density"])
            prev = row["l2_density"]
            rows.append(row)
            print(json.dumps(row), flush=True)
    print("\n| N | p | triangles | steps | L2 err (rho) FP64 | L2 err (rho) FP32-bulk | rate | FP32 vs FP64 rel. L2 | wall FP64 / FP32 (s) |")
    print("|---|---|---|---|---|---|---|---|---|")
    for r in rows:
        rate = "" if r["rate"] is None else f"{r['rate']:.2f}"
        print(f"| {r['n']} | {r['order']} | {r['triangles']} | {r['steps']} | {r['l2_density']:.3e} | "
              f"{r['l2_density_fp32']:.3e} | {rate} | {r['fp32_vs_fp64_rel_l2']:.2e} | {r['wall_s']:.1f} / {r['wall_fp32_s']:.1f} |")


if __name__ == "__main__":
    main()
