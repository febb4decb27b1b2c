"""The timing entry points (needs an MI355X).

crys_diffusion_time is the device counterpart of the reference's triplet
clock (std::clock around diffusion(), crysSCFT_sample.cpp:389-391, summed at
:512): mean milliseconds per diffusion() over back-to-back calls, from HIP
events on the solver's stream.  It must be positive and finite, lie within the
host's wall clock of the same calls, and agree with the per-launch times
crys_scft_time_step reports for the same kernels.
"""
import os
import subprocess
import sys
import time

import numpy as np
import pytest

from conftest import ROOT
import crysscft_sample_amd as C

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["PmmmDCT", "3m", "P1"])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_diffusion_time_against_the_wall_clock(name, prec):
    s = C.SOLVERS[name]((64, 64, 64), precision=prec)
    s.set_cell_para((4.0,) * 3)
    s.set_contour_step(0.01)
    s.upload_iomat(np.random.default_rng(5).random(s.get_grid_size_phy()))
    reps = 200
    ms = s.time_diffusion(reps)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        s.diffusion()
    s.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3 / reps
    print(f"  {name} {prec}: {ms * 1e3:.2f} us per diffusion() on the device, "
          f"{wall_ms * 1e3:.2f} us wall", flush=True, file=sys.stderr)
    assert np.isfinite(ms) and 0 < ms < 1.0
    assert ms <= 1.05 * wall_ms
    s.close()


STEP_CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import crysscft_sample_amd as C
from crysscft_sample_amd import fields as F
N = (64, 64, 64)
s = C.SolverPmmmDCT(N)
run = C.SCFT(s, 0.25, 0.005, 20.0, 4.0, 0.1)
wA, wB = F.bcc_fields(N, 4.0, 0.25, 20.0)
run.set_fields(F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT"))
ms_step, per = run.time_step(reps=100)
d = C.SolverPmmmDCT(N)
d.set_cell_para((4.0,) * 3)
d.set_contour_step(0.005)
d.upload_iomat(np.ones(d.get_grid_size_phy()))
print(json.dumps({"step": ms_step, "per_launch": per, "diffusion": d.time_diffusion(200)}))
"""


def test_diffusion_time_agrees_with_the_step_kernels():
    """With the one-launch phases off, a contour step of the SCFT iteration is
    the same fused z+y slab and x pass that one diffusion() launches, for two
    chains instead of one: the step's kernels take at least as long as one
    diffusion() and at most a few times it."""
    env = dict(os.environ, CRYS_PHASE="0")
    r = subprocess.run([sys.executable, "-c", STEP_CHILD, ROOT], env=env, capture_output=True,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    import json
    d = json.loads(r.stdout.strip().splitlines()[-1])
    print(f"  step {d['step'] * 1e3:.2f} us ({[round(v * 1e3, 2) for v in d['per_launch']]}), "
          f"diffusion {d['diffusion'] * 1e3:.2f} us", flush=True, file=sys.stderr)
    assert len(d["per_launch"]) == 2
    assert 0.8 * d["diffusion"] <= sum(d["per_launch"]) <= 4 * d["diffusion"]
    assert 0.8 * d["diffusion"] <= d["step"] <= 4 * d["diffusion"]


@pytest.mark.parametrize("workgroups", [64, 256])
def test_grid_sync_variants_exchange_correctly(workgroups):
    """Every synchronisation variant of crys_time_grid_sync delivers the
    exchanged values (none read before they were written) without a spin
    time-out, and reports a positive time per round."""
    import ctypes

    from crysscft_sample_amd._lib import check, lib

    for v in range(7):
        us = ctypes.c_double()
        bad, tmo = ctypes.c_uint(), ctypes.c_uint()
        check(lib().crys_time_grid_sync(v, workgroups, 50, ctypes.byref(us), ctypes.byref(bad),
                                        ctypes.byref(tmo)))
        print(f"  variant {v}, {workgroups} workgroups: {us.value:.2f} us per round", flush=True,
              file=sys.stderr)
        assert bad.value == 0 and tmo.value == 0
        assert 0 < us.value < 1e4
