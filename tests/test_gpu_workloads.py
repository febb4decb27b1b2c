"""The headline and batched workloads in full (needs an MI355X).

  - Headline: PmmmDCT on a 128^3 grid (64^3 reduced), fA 0.30, xN 25,
    L 4.2, ds 1/200 (NsA 60, NsB 140), BCC-seeded fields.  FP64 within 1e-10
    of the oracle in every trace column; FP32 propagators within
    |dF/F| <= 1e-7 of the oracle, with max |dphiA|, max |dphiB| stated.
  - FP32 on exactly uniform fields: E is one FP32 value whose rounding is
    applied 2 (NsA + NsB) times to every point alike, and the products round
    every point alike too, so neither averages out.  The error in ln Q is
    bounded by E's rounding (predicted exactly) plus one FP32 ulp per contour
    step of the backward chain, and E's rounding dominates it.
  - Batched: 256 independent PmmmDCT instances with reduced grids drawn from
    {16, 24, 32, 48, 64}^3, fA ~ U[0.15, 0.35], xN ~ U[15, 40],
    L ~ U[3.6, 4.6], Ns from {100, 150, 200}, BCC-seeded fields, on the
    autotuned path: FP32 against FP64 for every instance, and one instance per
    grid size and per Ns against the oracle (each instance reproduces its own
    diblock_scft_simple_iterate run, crysSCFT_sample.cpp:426-531).
"""
import sys

import numpy as np
import pytest

import crysscft_sample_amd as C
from crysscft_sample_amd import fields as F
from crysscft_sample_amd.batch import Batch, random_instances

pytestmark = pytest.mark.gpu

HEAD = dict(N=128, fA=0.30, xN=25.0, L=4.2, ds=1 / 200, acc=0.1)


def say(*a):
    print(*a, flush=True, file=sys.stderr)


@pytest.fixture(scope="module")
def headline_oracle(oracle):
    c = HEAD
    wA, wB = F.bcc_fields((c["N"],) * 3, c["L"], c["fA"], c["xN"])
    a, b = F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT")
    say("oracle: 2 iterations of the 128^3 headline workload")
    tr, wA2, wB2, phiA, phiB = oracle.Solver("PmmmDCT", (c["N"],) * 3).scft(
        c["fA"], c["ds"], c["xN"], c["L"], c["acc"], 2, a, b)
    return a, b, tr, phiA, phiB, wA2


def _gpu_headline(prec, a, b, nstep):
    c = HEAD
    s = C.SolverPmmmDCT((c["N"],) * 3, precision=prec)
    run = C.SCFT(s, c["fA"], c["ds"], c["xN"], c["L"], c["acc"])
    assert (run.NsA, run.NsB) == (60, 140)
    run.set_fields(a, b)
    tr = run.run(nstep)
    phiA, phiB, wA, _ = run.fields()
    run.close()
    s.close()
    return tr, phiA, phiB, wA


def test_headline_workload_fp64_against_oracle(headline_oracle):
    a, b, tr_ref, phiA_ref, phiB_ref, wA_ref = headline_oracle
    tr, phiA, phiB, wA = _gpu_headline("fp64", a, b, 2)
    col = np.abs(tr / tr_ref - 1).max(axis=0)
    say("  fp64 128^3 per-column max relative error:", " ".join(f"{v:.1e}" for v in col))
    assert col[0] <= 1e-10 and col[5] <= 1e-10
    assert (col[[1, 2, 3, 6]] <= 1e-10).all()
    assert np.abs(tr[:, 4] - tr_ref[:, 4]).max() <= 1e-10 * np.abs(tr_ref[:, 0]).max()
    assert np.abs(phiA - phiA_ref).max() <= 1e-10
    assert np.abs(phiB - phiB_ref).max() <= 1e-10
    assert np.abs(wA - wA_ref).max() <= 1e-10 * np.abs(wA_ref).max()


def test_headline_workload_fp32_against_oracle(headline_oracle):
    """FP32 propagators, FP32 accumulation and split-precision products; Q,
    phi, the energies and the mixing in FP64."""
    a, b, tr_ref, phiA_ref, phiB_ref, _ = headline_oracle
    tr, phiA, phiB, _ = _gpu_headline("fp32", a, b, 2)
    dF = np.abs(tr[:, 0] / tr_ref[:, 0] - 1).max()
    dlnQ = np.abs(tr[:, 2] - tr_ref[:, 2]).max()
    dM = np.abs(tr[:, 5] / tr_ref[:, 5] - 1).max()
    dpA = np.abs(phiA - phiA_ref).max()
    dpB = np.abs(phiB - phiB_ref).max()
    say(f"  fp32 128^3: |dF/F| {dF:.2e}  |d lnQ| {dlnQ:.2e}  |dMaxInc/MaxInc| {dM:.2e}  "
        f"max|dphiA| {dpA:.2e}  max|dphiB| {dpB:.2e}")
    assert dF <= 1e-7
    assert dpA <= 1e-6 and dpB <= 1e-6
    assert dM <= 1e-5


def test_fp32_uniform_field_error_is_the_rounding_of_E():
    c = HEAD
    a, b = 1.3, -0.7
    s = C.SolverPmmmDCT((c["N"],) * 3, precision="fp32")
    run = C.SCFT(s, c["fA"], c["ds"], c["xN"], c["L"], 0.0)
    shape = s.get_grid_size_phy()
    run.set_fields(np.full(shape, a), np.full(shape, b))
    tr = run.run(1)
    exact = c["ds"] * (a * run.NsA + b * run.NsB)  # -ln Q in closed form
    EA, EB = np.exp(-c["ds"] / 2 * a), np.exp(-c["ds"] / 2 * b)
    dA, dB = float(np.float32(EA)) / EA - 1, float(np.float32(EB)) / EB - 1
    predicted = -2 * (run.NsA * dA + run.NsB * dB)  # error of -ln Q from E alone
    got = tr[0, 2] - exact
    bound = abs(predicted) + (run.NsA + run.NsB) * 2.0 ** -24
    say(f"  uniform fp32: d(-lnQ) {got:.3e}, predicted from E's rounding {predicted:.3e}, "
        f"bound {bound:.3e}, |dF/F| {abs(got) / abs(tr[0, 0]):.2e}")
    assert abs(got) <= bound
    assert abs(got) >= 0.5 * abs(predicted) and np.sign(got) == np.sign(predicted)
    phiA = run.fields()[0]
    assert np.abs(phiA - c["fA"]).max() < 1e-6


@pytest.fixture(scope="module")
def batch256():
    insts = random_instances(256)
    ws = []
    for k, inst in enumerate(insts):
        wA, wB = F.bcc_fields(inst.grid_log, inst.cell, inst.fA, inst.xN, seed=1000 + k)
        ws.append((F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT")))
    out = {}
    for prec in ("fp64", "fp32"):
        b = Batch(insts, precision=prec)
        for k, (wA, wB) in enumerate(ws):
            b.set_fields(k, wA, wB)
        path, ms_ph, ms_pa = b.autotune(reps=1)
        say(f"  256 instances {prec}: autotune phases {ms_ph} ms, passes {ms_pa:.2f} ms -> {path}")
        tr = b.run(2)
        out[prec] = dict(trace=tr, path=path, phiA=[b.fields(k)[0] for k in range(len(insts))])
        b.close()
    return insts, ws, out


def test_batch_256_fp32_against_fp64(batch256):
    insts, _, out = batch256
    t64, t32 = out["fp64"]["trace"], out["fp32"]["trace"]
    assert t64.shape == (256, 2, 7) and np.isfinite(t64).all() and np.isfinite(t32).all()
    dF = np.abs(t32[:, :, 0] / t64[:, :, 0] - 1).max(axis=1)
    dphi = np.array([np.abs(a - b).max() for a, b in zip(out["fp32"]["phiA"],
                                                          out["fp64"]["phiA"])])
    say(f"  fp32 vs fp64, 256 instances: |dF/F| median {np.median(dF):.2e} max {dF.max():.2e}; "
        f"max|dphiA| median {np.median(dphi):.2e} max {dphi.max():.2e}")
    for k in np.argsort(dF)[-3:]:
        i = insts[k]
        say(f"    instance {k}: {i.grid_log[0] // 2}^3 Ns {round(1 / i.ds)} fA {i.fA:.3f} "
            f"xN {i.xN:.1f} L {i.cell[0]:.2f}: F {t64[k, -1, 0]:.6f} |dF/F| {dF[k]:.2e}")
    # measured: median 3.4e-8, max 1.2e-6
    assert np.median(dF) <= 1e-7
    assert dF.max() <= 3e-6
    assert dphi.max() <= 1e-5


def _sample(insts):
    """One instance per reduced grid size and per contour length."""
    pick = {}
    for k, inst in enumerate(insts):
        m, Ns = inst.grid_log[0] // 2, round(1 / inst.ds)
        pick.setdefault(("m", m), k)
        pick.setdefault(("Ns", Ns), k)
    return sorted(set(pick.values()))


def test_batch_256_instances_against_oracle(oracle, batch256):
    insts, ws, out = batch256
    sample = _sample(insts)
    sizes = {insts[k].grid_log[0] // 2 for k in sample}
    nss = {round(1 / insts[k].ds) for k in sample}
    assert sizes == {16, 24, 32, 48, 64} and nss == {100, 150, 200}
    for k in sample:
        inst = insts[k]
        t_ref, *_ = oracle.Solver("PmmmDCT", inst.grid_log).scft(
            inst.fA, inst.ds, inst.xN, inst.cell, inst.acceptance, 2, *ws[k])
        d64 = np.abs(out["fp64"]["trace"][k, :, 0] / t_ref[:, 0] - 1).max()
        m64 = np.abs(out["fp64"]["trace"][k, :, 5] / t_ref[:, 5] - 1).max()
        d32 = np.abs(out["fp32"]["trace"][k, :, 0] / t_ref[:, 0] - 1).max()
        say(f"  instance {k}: {inst.grid_log[0] // 2}^3 reduced, Ns {round(1 / inst.ds)}: "
            f"fp64 |dF/F| {d64:.1e} |dMaxInc| {m64:.1e}, fp32 |dF/F| {d32:.1e}")
        assert d64 <= 1e-10 and m64 <= 1e-10
        assert d32 <= 1e-6
