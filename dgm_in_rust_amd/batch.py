"""Many cases on one GPU: the dgm_batch_* entry points of libdgm.so.

A :class:`Batch` keeps one arena per polynomial order. Each case is a disconnected, 64-aligned piece
of its arena's mesh arrays, so every phase of a step is one kernel launch per order over all of its
cases, with per-case time steps, stop conditions, error status and lift. Cases are added and removed
between steps (continuous batching), and a case gives the same result alone as in any mix. The
reference runs one case per process and has no counterpart of this layer.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
from typing import Dict, List, Optional

import numpy as np

from . import _lib
from ._lib import DgmError, ptr
from .meshgen import MeshInput

_PREC = {"fp64": _lib.FP64, "fp32": _lib.FP32_BULK, "fp32_bulk": _lib.FP32_BULK}
_LIM = {"off": _lib.LIMITER_OFF, "jacobi": _lib.LIMITER_JACOBI, None: _lib.LIMITER_OFF}
INT32_MAX = 2**31 - 1


def n_basis(order: int) -> int:
    return (order + 1) * (order + 2) // 2


class Batch:
    """dgm_batch: ``capacity`` maps order p -> element capacity of its arena."""

    def __init__(self, capacity: Dict[int, int], precision: str = "fp32", limiter: Optional[str] = "jacobi",
                 hcr: float = 1.4, max_cases: int = 256):
        L = _lib.lib()
        self._L = L
        cap = (C.c_int64 * 4)(*[int(capacity.get(p, 0)) for p in (1, 2, 3, 4)])
        self.params = _lib.BatchParams(_PREC[precision], _LIM[limiter], hcr, max_cases, 0, cap)
        self.precision, self.limiter, self.hcr, self.max_cases = precision, limiter, hcr, max_cases
        h = C.c_void_p()
        s = L.dgm_batch_create(C.byref(self.params), C.byref(h))
        if s != 0:
            raise DgmError(s, message="dgm_batch_create failed")
        self.h = h
        self._case_order: Dict[int, int] = {}
        self._case_ne: Dict[int, int] = {}

    # ---- lifetime ----
    def close(self):
        if getattr(self, "h", None):
            self._L.dgm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, s, what):
        if s != 0:
            e = _lib.Error()
            self._L.dgm_batch_last_error(self.h, C.byref(e))
            raise DgmError(s, e.edge, e.element, f"{what}: {e.message.decode(errors='replace')}")

    # ---- cases ----
    def footprint(self, order: int, n_elements: int, n_edges: int, n_patches: int) -> int:
        b = C.c_int64()
        self._check(self._L.dgm_batch_footprint(self.h, order, n_elements, n_edges, n_patches, C.byref(b)),
                    "dgm_batch_footprint")
        return int(b.value)

    def add(self, mesh: MeshInput, order: int, mach: float, aoa_deg: float, rho: float = 1.0, p: float = 1.0e5,
            fs_hcr: float = 1.4, cfl: float = 0.0, final_step: int = INT32_MAX, final_time: float = 1.0e300,
            U0: Optional[np.ndarray] = None) -> Optional[int]:
        """Admits a case and returns its slot, or None when its arena has no free range now (NO_SPACE)."""
        m = mesh.contiguous()
        ms = _lib.mesh_struct(m)
        cp = _lib.CaseParams(order, int(min(final_step, INT32_MAX)), final_time, cfl, mach, aoa_deg, rho, p, fs_hcr)
        u = None if U0 is None else np.ascontiguousarray(U0, np.float64)
        cid = C.c_int32(-1)
        s = self._L.dgm_batch_add(self.h, C.byref(ms), C.byref(cp), ptr(u), C.byref(cid))
        if s == _lib.NO_SPACE:
            return None
        self._check(s, "dgm_batch_add")
        i = int(cid.value)
        self._case_order[i] = order
        self._case_ne[i] = int(sum(int(c) for c in m.el_count))
        return i

    def remove(self, case_id: int):
        """Retires a case (stream-ordered, without waiting for the queued steps)."""
        self._check(self._L.dgm_batch_remove(self.h, case_id), "dgm_batch_remove")
        self._case_order.pop(case_id, None)
        self._case_ne.pop(case_id, None)

    def steps_async(self, n: int):
        self._check(self._L.dgm_batch_steps_async(self.h, n), "dgm_batch_steps_async")

    def synchronize(self):
        self._check(self._L.dgm_batch_synchronize(self.h), "dgm_batch_synchronize")

    def info(self, n: Optional[int] = None) -> List[dict]:
        """Control records of slots 0 .. n-1 as dicts (live, active, status, steps, time, ...)."""
        n = self.max_cases if n is None else n
        arr = (_lib.CaseInfo * n)()
        self._check(self._L.dgm_batch_case_info(self.h, n, arr), "dgm_batch_case_info")
        return [{k: getattr(a, k) for k, _ in _lib.CaseInfo._fields_} for a in arr]

    def state(self, case_id: int) -> np.ndarray:
        U = np.zeros((self._case_ne[case_id], 4, n_basis(self._case_order[case_id])))
        self._check(self._L.dgm_batch_case_state(self.h, case_id, ptr(U)), "dgm_batch_case_state")
        return U

    # ---- serving polls ----
    def poll_async(self, chord: float = 1.0, cl_tol: float = 1e-5) -> int:
        """Queues a poll after the work already queued (see dgm_batch_poll_async); returns its ticket."""
        t = C.c_int64(-1)
        self._check(self._L.dgm_batch_poll_async(self.h, chord, cl_tol, C.byref(t)), "dgm_batch_poll_async")
        return int(t.value)

    def poll_wait(self, ticket: int, n: Optional[int] = None):
        """Waits for the poll `ticket` (the oldest outstanding); returns (records as dicts, lift coefficients,
        converged marks) of slots 0 .. n-1 as they were when the poll ran."""
        n = self.max_cases if n is None else n
        arr = (_lib.CaseInfo * n)()
        cl = np.zeros(n)
        conv = np.zeros(n, np.int32)
        self._check(self._L.dgm_batch_poll_wait(self.h, ticket, n, arr, ptr(cl), ptr(conv, _lib._i32p)),
                    "dgm_batch_poll_wait")
        return [{k: getattr(a, k) for k, _ in _lib.CaseInfo._fields_} for a in arr], cl, conv.astype(bool)

    def device_ms(self) -> float:
        """Device time of the completed step launches so far, in ms (events around each steps_async call)."""
        v = C.c_double(0.0)
        self._check(self._L.dgm_batch_device_time(self.h, C.byref(v)), "dgm_batch_device_time")
        return float(v.value)

    def lift(self, chord: float = 1.0, n: Optional[int] = None) -> np.ndarray:
        n = self.max_cases if n is None else n
        cl = np.zeros(n)
        self._check(self._L.dgm_batch_lift(self.h, chord, n, ptr(cl)), "dgm_batch_lift")
        return cl

    # ---- launch shapes ----
    def tuning(self, order: int) -> dict:
        t = _lib.Tuning()
        self._check(self._L.dgm_batch_get_tuning(self.h, order, C.byref(t)), "dgm_batch_get_tuning")
        return t.as_dict()

    def set_tuning(self, order: int, values: dict):
        t = _lib.Tuning(**{k: int(values[k]) for k, _ in _lib.Tuning._fields_})
        self._check(self._L.dgm_batch_set_tuning(self.h, order, C.byref(t)), "dgm_batch_set_tuning")

    def profile_steps(self, n: int = 5) -> Dict[int, dict]:
        """Mean ms per launch of each order's kernels over n eager steps, the state and clocks restored
        afterwards (dgm_batch_profile_steps): {order: {surface_ms, element_ms, detect_ms, rebuild_ms,
        apply_ms, step_ms}} for the orders with resident cases."""
        out = (_lib.KernelTimes * 4)()
        self._check(self._L.dgm_batch_profile_steps(self.h, n, out), "dgm_batch_profile_steps")
        keys = ("surface_ms", "element_ms", "detect_ms", "rebuild_ms", "apply_ms", "step_ms")
        return {p: {k: float(getattr(out[p - 1], k)) for k in keys} for p in sorted(set(self._case_order.values()))}

    def autotune(self, repeats: int = 10, cache: Optional[str] = None) -> Dict[int, dict]:
        """Times the launch shapes of every order with resident cases and keeps the fastest (see
        dgm_batch_autotune). With ``cache`` (a JSON file), a result found for the same order, precision,
        limiter and arena size bucket (a power of two of the resident elements) is reused instead."""
        orders = sorted(set(self._case_order.values()))
        key = {}
        for p in orders:
            ne = sum(self._case_ne[c] for c, o in self._case_order.items() if o == p)
            key[p] = f"p{p}/{self.precision}/{self.limiter}/2^{int(math.log2(max(ne, 1)))}"
        table = {}
        if cache and os.path.exists(cache):
            with open(cache) as f:
                table = json.load(f)
        if cache and orders and all(key[p] in table for p in orders):
            for p in orders:
                self.set_tuning(p, table[key[p]])
            return {p: table[key[p]] for p in orders}
        out = (_lib.Tuning * 4)()
        self._check(self._L.dgm_batch_autotune(self.h, repeats, out), "dgm_batch_autotune")
        chosen = {p: out[p - 1].as_dict() for p in orders}
        if cache:
            table.update({key[p]: chosen[p] for p in orders})
            d = os.path.dirname(os.path.abspath(cache))
            os.makedirs(d, exist_ok=True)
            with open(cache, "w") as f:
                json.dump(table, f, indent=1, sort_keys=True)
        return chosen
