"""GPU: the launch-boundary study's kernels (libdgm_study.so, outside the library) run their passes and
report positive times; the study checks the buffer itself (every float counts every pass of every
repetition)."""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "dgm_in_rust_amd", "libdgm_study.so")


def test_sync_study_counts_every_pass():
    lib = C.CDLL(LIB)
    lib.dgm_study_sync.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    maxg = lib.dgm_study_max_grid(256)
    assert maxg >= 256
    g, c = C.c_double(), C.c_double()
    assert lib.dgm_study_sync(256, 256, 3, 4, C.byref(g), C.byref(c)) == 0
    assert g.value > 0 and c.value > 0
    # a grid larger than can be co-resident is refused before any cooperative launch
    assert lib.dgm_study_sync(maxg + 8, 256, 3, 1, C.byref(g), C.byref(c)) == 1
