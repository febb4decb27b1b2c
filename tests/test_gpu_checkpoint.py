"""Contour checkpointing of the SCFT iteration and of crys_scft_evaluate (needs an MI355X).

With checkpoints every C slots the forward chain is kept only at the slots
0, C, 2C, ...; each segment between two checkpoints is recomputed while the
backward chain advances over it, and the contour sums of phi and the stress
are taken segment by segment.  Checks:
  - FP64 SCFT traces with checkpoints against the oracle to 1e-10 on the
    phase path (class 32 centrosymmetric, class 48) and on the per-axis pass
    path, for intervals 1, 7 (not a divisor of Ns), 12 and the automatic one;
  - the same runs equal the full-history GPU run to 1e-12 (the propagators
    are recomputed bit for bit; only the order of the contour sums differs);
  - crys_scft_evaluate every 12 slots on the differentiation workload (64^3
    logical, orthorhombic (4.0, 4.1, 4.3), fA = 0.28, xN = 22, Ns = 35 + 92):
    H, dH/dw and dH/dL equal the full-history values, and dH/dL matches
    fourth-order central differences of H to 1e-6;
  - the automatic choice: a budget below the full histories selects
    ceil(sqrt(Ns)) and stays within the budget; a budget below any
    checkpointed storage fails with CRYS_ERR_OUT_OF_MEMORY;
  - FP32 propagators with checkpoints agree with FP32 full histories to
    1e-6 (measured 1.5e-8 in the trace on the class-32 phase path, where a
    segment restarts from the stored FP32 history value and the multiply by
    the next E may contract differently; 5e-17 on the per-axis passes);
  - repeated runs are bitwise identical (deterministic reductions), with and
    without checkpoints.
"""
import math
import sys

import numpy as np
import pytest

from conftest import rel_err
import crysscft_sample_amd as C
from crysscft_sample_amd import _lib
from crysscft_sample_amd import fields as F

pytestmark = pytest.mark.gpu

FA, XN, DS, ACC, NSTEP = 0.3, 25.0, 0.01, 0.1, 3

CASES = [
    ("PmmmDCT", (64, 64, 64), (4.2,) * 3),   # phases, class 32, centrosymmetric
    ("P1", (48, 48, 48), (4.0,) * 3),        # phases, class 48
    ("PmmmDCT", (40, 64, 96), (4.1,) * 3),   # reduced 20 x 32 x 48: per-axis passes
]


def say(*a):
    print(*a, flush=True, file=sys.stderr)


def run_gpu(name, N, cell, a, b, every=None, budget=0, precision="fp64"):
    s = C.SOLVERS[name](N, precision=precision)
    run = C.SCFT(s, FA, DS, XN, cell, ACC)
    if every is not None:
        run.set_checkpointing(every, budget)
    used = run.checkpointing()
    run.set_fields(a, b)
    tr = run.run(NSTEP)
    phiA, phiB, wA, wB = run.fields()
    run.close()
    s.close()
    return tr, phiA, wA, used


@pytest.fixture(scope="module")
def references(oracle):
    out = {}
    for name, N, cell in CASES:
        wA, wB = F.bcc_fields(N, cell, FA, XN)
        a, b = F.extract(wA, name), F.extract(wB, name)
        tr, _, _, phiA, _ = oracle.Solver(name, N).scft(FA, DS, XN, cell, ACC, NSTEP, a, b)
        full = run_gpu(name, N, cell, a, b)
        assert full[3][0] == 0
        out[name, N] = (a, b, tr, phiA, full)
    return out


@pytest.mark.parametrize("every", [1, 7, 12, "auto"])
@pytest.mark.parametrize("name,N,cell", CASES)
def test_checkpointed_iteration(references, name, N, cell, every):
    a, b, tr, phiA, (tr_full, phi_full, w_full, _) = references[name, N]
    Ns = int(FA / DS) + int((1 - FA) / DS)  # the truncation of crysSCFT_sample.cpp:443-444
    if every == "auto":
        # a budget of a third of the full histories forces checkpointing
        budget = 2 * (Ns + 1) * a.size * 8 // 3
        got, gphi, gw, (used, nbytes) = run_gpu(name, N, cell, a, b, 0, budget)
        assert used == math.ceil(math.sqrt(Ns))
        assert nbytes <= budget
    else:
        got, gphi, gw, (used, nbytes) = run_gpu(name, N, cell, a, b, every)
        assert used == every
    dF = np.abs(got[:, 0] / tr[:, 0] - 1).max()
    dM = np.abs(got[:, 5] / tr[:, 5] - 1).max()
    say(f"checkpoint {name} {N} every {used}: |dF/F| {dF:.2e} |dMaxInc| {dM:.2e} "
        f"vs full {rel_err(got, tr_full):.2e}, storage {nbytes / 2**20:.1f} MiB")
    assert dF <= 1e-10 and dM <= 1e-10
    assert rel_err(gphi, phiA) < 1e-10
    assert rel_err(got, tr_full) < 1e-12
    assert rel_err(gphi, phi_full) < 1e-12
    assert rel_err(gw, w_full) < 1e-12


def test_automatic_interval_and_budget():
    name, N, cell = CASES[0]
    s = C.SOLVERS[name](N)
    run = C.SCFT(s, FA, DS, XN, cell, ACC)
    Ns = run.NsA + run.NsB
    n = int(np.prod(s.get_grid_size_phy()))
    full_bytes = 2 * (Ns + 1) * n * 8
    run.set_checkpointing(0, 0)  # free memory holds the full histories
    assert run.checkpointing() == (0, full_bytes)
    run.set_checkpointing(0, full_bytes - 1)
    every, nbytes = run.checkpointing()
    assert every == math.ceil(math.sqrt(Ns))
    nck = -(-Ns // every)
    assert nbytes == (nck + 2 * (every + 1)) * n * 8 < full_bytes
    with pytest.raises(C.CrysError) as e:
        run.set_checkpointing(0, 10 * n * 8)
    assert e.value.status == _lib.ERR_OUT_OF_MEMORY
    assert run.checkpointing() == (every, nbytes)  # the previous setting stays
    with pytest.raises(C.CrysError) as e:
        run.set_checkpointing(12, 10 * n * 8)
    assert e.value.status == _lib.ERR_OUT_OF_MEMORY
    # a usable setting again afterwards
    run.set_checkpointing(5, 0)
    assert run.checkpointing()[0] == 5
    run.close()
    s.close()


def test_fp32_checkpointed_equals_full():
    name, N, cell = CASES[0]
    wA, wB = F.bcc_fields(N, cell, FA, XN)
    a, b = F.extract(wA, name), F.extract(wB, name)
    full = run_gpu(name, N, cell, a, b, precision="fp32")
    ck = run_gpu(name, N, cell, a, b, 9, precision="fp32")
    say(f"fp32 checkpointed vs full: trace {rel_err(ck[0], full[0]):.2e} "
        f"phiA {rel_err(ck[1], full[1]):.2e}")
    assert rel_err(ck[0], full[0]) < 1e-7  # measured 2.3e-8
    assert rel_err(ck[1], full[1]) < 5e-7  # measured 9.3e-8


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("every", [None, 9])
def test_runs_are_bitwise_repeatable(prec, every):
    name, N, cell = CASES[0]
    wA, wB = F.bcc_fields(N, cell, FA, XN)
    a, b = F.extract(wA, name), F.extract(wB, name)
    r1 = run_gpu(name, N, cell, a, b, every, precision=prec)
    r2 = run_gpu(name, N, cell, a, b, every, precision=prec)
    for x, y in zip(r1[:3], r2[:3]):
        assert np.array_equal(x, y)


# Differentiation workload: 64^3 logical, orthorhombic cell, Ns = 35 + 92.
DN = (64, 64, 64)
DCELL = (4.0, 4.1, 4.3)
DFA, DXN, DDS = 0.28, 22.0, 1.0 / 128


def central4(f, h):
    return (-f(2 * h) + 8 * f(h) - 8 * f(-h) + f(-2 * h)) / (12 * h)


def test_evaluate_every_12_slots():
    wAf, wBf = F.bcc_fields(DN, DCELL, DFA, DXN)
    a, b = F.extract(wAf, "PmmmDCT"), F.extract(wBf, "PmmmDCT")
    s = C.SolverPmmmDCT(DN, orthorhombic=True)
    run = C.SCFT(s, DFA, DDS, DXN, DCELL, 0.0)
    assert (run.NsA, run.NsB) == (35, 92)
    g_full = run.evaluate(a, b, stress=True)
    run.set_checkpointing(12)
    assert run.checkpointing()[0] == 12
    g = run.evaluate(a, b, stress=True)
    say(f"evaluate every 12: H {abs(g['H'] / g_full['H'] - 1):.2e}, "
        f"dH/dwA {rel_err(g['dH_dwA'], g_full['dH_dwA']):.2e}, "
        f"dH/dL {rel_err(g['dH_dL'], g_full['dH_dL']):.2e}")
    assert abs(g["H"] / g_full["H"] - 1) < 1e-12
    assert rel_err(g["dH_dwA"], g_full["dH_dwA"]) < 1e-12
    assert rel_err(g["dH_dwB"], g_full["dH_dwB"]) < 1e-12
    assert rel_err(g["dH_dL"], g_full["dH_dL"]) < 1e-12
    h = 1e-4
    for ax in range(3):
        def H(d):
            cell = list(DCELL)
            cell[ax] += d
            s.set_cell_para(cell)
            s.set_contour_step(DDS)
            return run.evaluate(a, b, stress=False)["H"]
        fd = central4(H, h)
        s.set_cell_para(DCELL)
        s.set_contour_step(DDS)
        say(f"  dH/dL{ax}: fd {fd:.10e} checkpointed adjoint {g['dH_dL'][ax]:.10e}")
        assert abs(fd / g["dH_dL"][ax] - 1) < 1e-6
    run.close()
    s.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_differentiation_workload_gradient_against_oracle(oracle, prec):
    """The differentiation workload at its full size (64^3, orthorhombic,
    Ns 35 + 92, checkpoints every 12 slots): dH/dw against the oracle's own
    densities, d(-ln Q)/dwA = (ds NsA / fA) phiA / n (the oracle's P1 on the
    full mirror-symmetric fields, which honours the orthorhombic cell), and
    against central differences of H at a few points."""
    wAf, wBf = F.bcc_fields(DN, DCELL, DFA, DXN)
    a, b = F.extract(wAf, "PmmmDCT"), F.extract(wBf, "PmmmDCT")
    s = C.SolverPmmmDCT(DN, orthorhombic=True, precision=prec)
    run = C.SCFT(s, DFA, DDS, DXN, DCELL, 0.0)
    run.set_checkpointing(12)
    g = run.evaluate(a, b, stress=False)
    tr, _, _, phiA, phiB = oracle.Solver("P1", DN).scft(DFA, DDS, DXN, DCELL, 0.0, 1, wAf, wBf)
    phiA, phiB = F.extract(phiA, "PmmmDCT"), F.extract(phiB, "PmmmDCT")
    n = a.size
    cA, cB = DDS * 35 / DFA, DDS * 92 / (1 - DFA)
    gA = ((a - b) / (2 * DXN) - 0.5 + cA * phiA) / n
    gB = (-(a - b) / (2 * DXN) - 0.5 + cB * phiB) / n
    H = float(np.mean((wAf - wBf) ** 2 / (4 * DXN) - (wAf + wBf) / 2 + DXN / 4) - np.log(tr[0, 6]))
    eA, eB = rel_err(g["dH_dwA"], gA), rel_err(g["dH_dwB"], gB)
    say(f"  {prec} 64^3 every 12: dH/dwA {eA:.2e}, dH/dwB {eB:.2e}, H {abs(g['H'] / H - 1):.2e}")
    tol = 1e-10 if prec == "fp64" else 1e-5
    assert eA < tol and eB < tol
    assert abs(g["H"] / H - 1) < (1e-11 if prec == "fp64" else 1e-6)
    if prec == "fp64":
        h = 2e-3
        for idx in [(0, 0, 0), (11, 23, 5), (31, 2, 17)]:
            def Hd(d):
                w = a.copy()
                w[idx] += d
                return run.evaluate(w, b, stress=False)["H"]
            fd = central4(Hd, h)
            say(f"    dH/dwA{idx}: fd {fd:.10e} adjoint {g['dH_dwA'][idx]:.10e}")
            assert abs(fd / g["dH_dwA"][idx] - 1) < 1e-6
    run.close()
    s.close()


def test_glide_3m_checkpointed_equals_full():
    """3m with translational parts (run as P1 on the expanded fields) under
    checkpointing: iteration and gradient equal the full-history run."""
    N, cell, g = (32, 32, 32), (4.0,) * 3, [0.5, 0.5, 0, 0, 0, 0, 0, 0, 0]
    red = 0.3 * np.random.default_rng(23).standard_normal((16, 16, 16))
    wA, wB = 20.0 * red, -20.0 * red
    out = {}
    for every in (None, 7):
        s = C.Solver3m(N, translational_part=g)
        run = C.SCFT(s, FA, DS, 20.0, cell, ACC)
        if every:
            run.set_checkpointing(every)
        ev = run.evaluate(wA, wB, stress=True)
        run.set_fields(wA, wB)
        out[every] = (run.run(3), ev)
        run.close()
        s.close()
    (tr0, ev0), (tr1, ev1) = out[None], out[7]
    say(f"glide 3m every 7: trace {rel_err(tr1, tr0):.2e}, grad {rel_err(ev1['dH_dwA'], ev0['dH_dwA']):.2e}")
    assert rel_err(tr1, tr0) < 1e-12
    assert rel_err(ev1["dH_dwA"], ev0["dH_dwA"]) < 1e-12
    assert rel_err(ev1["dH_dL"], ev0["dH_dL"]) < 1e-12


def test_fp32_evaluate_checkpointed():
    wAf, wBf = F.bcc_fields(DN, DCELL, DFA, DXN)
    a, b = F.extract(wAf, "PmmmDCT"), F.extract(wBf, "PmmmDCT")
    res = {}
    for every in (None, 12):
        s = C.SolverPmmmDCT(DN, precision="fp32", orthorhombic=True)
        run = C.SCFT(s, DFA, DDS, DXN, DCELL, 0.0)
        if every:
            run.set_checkpointing(every)
        res[every] = run.evaluate(a, b, stress=True)
        run.close()
        s.close()
    g0, g1 = res[None], res[12]
    say(f"fp32 evaluate every 12: H {abs(g1['H'] / g0['H'] - 1):.2e} "
        f"grad {rel_err(g1['dH_dwA'], g0['dH_dwA']):.2e} stress {rel_err(g1['dH_dL'], g0['dH_dL']):.2e}")
    # measured 2.6e-9, 7.4e-7 and 6.3e-8
    assert abs(g1["H"] / g0["H"] - 1) < 2e-8
    assert rel_err(g1["dH_dwA"], g0["dH_dwA"]) < 3e-6
    assert rel_err(g1["dH_dL"], g0["dH_dL"]) < 5e-7
