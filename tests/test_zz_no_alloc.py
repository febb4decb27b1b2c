"""crys_propagate_noalloc against crys_propagate (needs an MI355X).

The variant that allocates nothing stages the half-finished contour step in
the library's static device field instead of the solver's scratch field; its
histories and final iomat must equal crys_propagate's bit for bit, on two
calls in a row with the second one smaller (the static field keeps the first
call's contents), and the call must leave the free device memory unchanged.
"""
import ctypes
import sys

import numpy as np
import pytest

import crysscft_sample_amd as C
from crysscft_sample_amd import _lib

pytestmark = pytest.mark.gpu


def _hip():
    return ctypes.CDLL("libamdhip64.so.7")


def _free_bytes(hip):
    free, total = ctypes.c_size_t(), ctypes.c_size_t()
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _both(solver, Ns, direction, seed):
    """(q, iomat) from crys_propagate and from crys_propagate_noalloc on the
    same device inputs, and the free-memory drop across the no-alloc call."""
    hip = _hip()
    n = solver.get_space_size_phy()
    rng = np.random.default_rng(seed)
    E = np.exp(-0.005 * rng.standard_normal(n))
    init = 1.0 + 0.1 * rng.standard_normal(n)
    d = ctypes.c_void_p()
    nb = (2 + 2 * (Ns + 1)) * n * 8
    assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(nb)) == 0
    try:
        dE, dI = d.value, d.value + n * 8
        dq = [d.value + 2 * n * 8, d.value + (2 + Ns + 1) * n * 8]
        hip.hipMemcpy(ctypes.c_void_p(dE), E.ctypes.data_as(ctypes.c_void_p),
                      ctypes.c_size_t(n * 8), 1)
        hip.hipMemcpy(ctypes.c_void_p(dI), init.ctypes.data_as(ctypes.c_void_p),
                      ctypes.c_size_t(n * 8), 1)
        out = []
        drop = None
        for k, fn in enumerate((_lib.lib().crys_propagate, _lib.lib().crys_propagate_noalloc)):
            before = _free_bytes(hip)
            _lib.check(fn(solver.handle, dE, dI, Ns, direction, dq[k]))
            solver.synchronize()
            if k == 1:
                drop = before - _free_bytes(hip)
            q = np.empty((Ns + 1, n))
            hip.hipMemcpy(q.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dq[k]),
                          ctypes.c_size_t((Ns + 1) * n * 8), 2)
            out.append((q, solver.download_iomat().reshape(-1)))
    finally:
        hip.hipFree(d)
    return out, drop


CASES = [  # the largest physical grid (2^21 points, the whole static field), then smaller
    ("P1", (128, 128, 128), None, 4, 1),
    ("PmmmDCT", (64, 48, 32), None, 6, -1),
    ("3m", (32, 32, 32), [0.5, 0.5, 0, 0, 0, 0, 0, 0, 0], 3, 1),
]


def test_noalloc_propagate_equals_propagate():
    for k, (name, N, g, Ns, direction) in enumerate(CASES):
        print(f"noalloc case {k}: {name} {N} Ns {Ns} direction {direction}", flush=True,
              file=sys.stderr)
        s = (C.Solver3m(N, translational_part=g) if g else C.SOLVERS[name](N))
        s.set_cell_para((4.1, 4.2, 4.3))
        s.set_contour_step(0.01)
        (a, b), drop = _both(s, Ns, direction, 30 + k)
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1], b[1])
        assert np.isfinite(a[0]).all()
        print(f"  equal; free device memory drop across the call: {drop} bytes", flush=True,
              file=sys.stderr)
        assert drop <= 0
        s.close()


def test_noalloc_rejects_bad_arguments():
    s = C.SolverPmmmDCT((32, 32, 32))
    s.set_contour_step(0.01)
    with pytest.raises(C.CrysError) as e:
        _lib.check(_lib.lib().crys_propagate_noalloc(s.handle, None, None, 3, 1, None))
    assert e.value.status == _lib.ERR_INVALID_ARGUMENT
    s.close()
