"""Pins the CPU oracle (CPU only).

The oracle restates the reference (crysSCFT_sample.cpp, solver_*.h) with its
own transforms; here it is checked against scipy/numpy golden fixtures,
brute-force definitions, closed forms and the reference's own acceptance
criterion (identical traces from the three solvers, Readme.md:28-76), each
chosen so that a dropped term, a wrong sign or index, or a swapped axis
fails at least one of them.
"""
import numpy as np
import pytest

from conftest import CFG0, rel_err
from crysscft_sample_amd import fields as F


def brute_dct2_1d(x):
    n = len(x)
    j = np.arange(n)
    return np.array([2 * np.sum(x * np.cos(np.pi * (j + 0.5) * k / n)) for k in range(n)])


def brute_dct3_1d(x):
    n = len(x)
    j = np.arange(1, n)
    return np.array([x[0] + 2 * np.sum(x[1:] * np.cos(np.pi * j * (k + 0.5) / n))
                     for k in range(n)])


def apply_axes(fn, a):
    for ax in range(3):
        a = np.apply_along_axis(fn, ax, a)
    return a


# ---- transforms -----------------------------------------------------------
@pytest.mark.parametrize("i", range(4))
def test_transforms_match_golden(oracle, golden, i):
    g = golden["transforms"]
    a = g[f"in{i}"]
    assert rel_err(oracle.dct2_3d(a), g[f"dct2_{i}"]) < 1e-14
    assert rel_err(oracle.dct3_3d(a), g[f"dct3_{i}"]) < 1e-14
    assert rel_err(oracle.rfft_3d(a), g[f"rfft_{i}"]) < 1e-14
    c = g[f"cin{i}"]
    assert rel_err(oracle.irfft_3d(c, a.shape[2]), g[f"irfft_{i}"]) < 1e-14


@pytest.mark.parametrize("shape", [(2, 3, 4), (5, 4, 3), (1, 6, 8)])
def test_dct_brute_force(oracle, shape):
    a = np.random.default_rng(7).standard_normal(shape)
    assert rel_err(oracle.dct2_3d(a), apply_axes(brute_dct2_1d, a)) < 1e-13
    assert rel_err(oracle.dct3_3d(a), apply_axes(brute_dct3_1d, a)) < 1e-13


def test_dct_pair_scale(oracle):
    # REDFT01(REDFT10(x)) = 2n x per axis (FFTW's documented normalisation)
    a = np.random.default_rng(8).standard_normal((4, 6, 8))
    assert rel_err(oracle.dct3_3d(oracle.dct2_3d(a)), a * 8 * 12 * 16) < 1e-14


# ---- solvers ----------------------------------------------------------------
@pytest.mark.parametrize("i", range(3))
@pytest.mark.parametrize("name,key", [("P1", "p1"), ("PmmmDCT", "pmmm"), ("3m", "m3")])
def test_diffusion_matches_golden(oracle, golden, i, name, key):
    d = golden["diffusion"]
    s = oracle.Solver(name, tuple(d[f"grid{i}"]))
    s.set_cell_para(tuple(d[f"cell{i}"]))
    s.set_contour_step(float(d[f"ds{i}"]))
    assert rel_err(s.apply(d[f"{key}_in{i}"]), d[f"{key}_out{i}"]) < 1e-13


def sym_full(shape, seed):
    f = np.random.default_rng(seed).standard_normal(shape)
    return F.mirror_symmetrise(f)


@pytest.mark.parametrize("N,cell", [((8, 12, 16), (4.0, 4.1, 4.3)), ((16, 16, 16), (4.2,) * 3)])
def test_cross_solver_identity(oracle, N, cell):
    f = sym_full(N, 11)
    p1 = oracle.Solver("P1", N)
    p1.set_cell_para(cell)
    p1.set_contour_step(0.01)
    r1 = p1.apply(f)
    m3 = oracle.Solver("3m", N)
    m3.set_cell_para(cell)
    m3.set_contour_step(0.01)
    assert rel_err(m3.apply(F.extract(f, "3m")), F.extract(r1, "3m")) < 1e-14
    if cell[0] == cell[1] == cell[2]:
        pm = oracle.Solver("PmmmDCT", N)
        pm.set_cell_para(cell)
        pm.set_contour_step(0.01)
        assert rel_err(pm.apply(F.extract(f, "PmmmDCT")), F.extract(r1, "PmmmDCT")) < 1e-14


def test_pmmm_plane_wave_eigenvector(oracle):
    # cos(pi k (j+1/2)/M) per axis is an eigenvector with eigenvalue K_k
    N, L, ds = (16, 12, 16), 4.3, 0.02
    M = [n // 2 for n in N]
    k = (3, 2, 5)
    vec = [np.cos(np.pi * k[a] * (np.arange(M[a]) + 0.5) / M[a]) for a in range(3)]
    f = vec[0][:, None, None] * vec[1][None, :, None] * vec[2][None, None, :]
    s = oracle.Solver("PmmmDCT", N)
    s.set_cell_para((L, L, L))
    s.set_contour_step(ds)
    lam = np.exp(-ds * sum((2 * np.pi * kk / L) ** 2 for kk in k))
    assert rel_err(s.apply(f), lam * f) < 1e-13


def test_p1_plane_wave_eigenvector(oracle):
    N, cell, ds = (8, 12, 16), (4.0, 4.1, 4.3), 0.02
    k = (3, -2, 5)
    x = [np.arange(n) for n in N]
    ph = (2 * np.pi * (k[0] * x[0][:, None, None] / N[0] + k[1] * x[1][None, :, None] / N[1]
                       + k[2] * x[2][None, None, :] / N[2]))
    f = np.cos(ph) + 0.5 * np.sin(ph)
    s = oracle.Solver("P1", N)
    s.set_cell_para(cell)
    s.set_contour_step(ds)
    lam = np.exp(-ds * sum((2 * np.pi * k[a] / cell[a]) ** 2 for a in range(3)))
    assert rel_err(s.apply(f), lam * f) < 1e-13


def test_pmmm_cubic_quirk_and_early_return(oracle):
    s = oracle.Solver("PmmmDCT", (8, 8, 16))
    np.testing.assert_array_equal(s.cell_para, [4, 4, 4, 0, 0, 0])  # solver_base.h:43
    s.set_cell_para((4.0, 4.1, 4.3))
    np.testing.assert_array_equal(s.cell_para, [4, 4, 4, 0, 0, 0])  # solver_Pmmm.h:101
    p1 = oracle.Solver("P1", (8, 8, 16))
    p1.set_cell_para((4.0, 4.1, 4.3))
    np.testing.assert_array_equal(p1.cell_para, [4.0, 4.1, 4.3, 0, 0, 0])


def test_bad_grids_rejected(oracle):
    with pytest.raises(ValueError):
        oracle.Solver("P1", (8, 8, 12))  # Nz % 8
    with pytest.raises(ValueError):
        oracle.Solver("3m", (8, 9, 16))  # odd
    with pytest.raises(ValueError):
        oracle.Solver("PmmmDCT", (8, 8, 24))  # (Nz/2) % 8


# ---- driver -------------------------------------------------------------
@pytest.mark.parametrize("name", ["P1", "3m", "PmmmDCT"])
def test_uniform_field_closed_form(oracle, name):
    N = (8, 8, 16)
    fA, ds, xN, a, b = 0.3, 0.01, 12.0, 0.7, -0.4
    s = oracle.Solver(name, N)
    shape = s.grid_size_phy
    tr, wA, wB, phiA, phiB = s.scft(fA, ds, xN, 4.0, 0.0, 1, np.full(shape, a), np.full(shape, b))
    NsA, NsB = int(fA / ds), int((1 - fA) / ds)
    Q = np.exp(-ds * (a * NsA + b * NsB))
    assert abs(tr[0, 6] / Q - 1) < 1e-13
    assert np.abs(phiA - fA).max() < 1e-13
    assert np.abs(phiB - (1 - fA)).max() < 1e-13


def test_mean_phi_equals_block_fraction(oracle):
    N = 16
    fA, xN, L, ds = 0.28, 22.0, 4.0, 1 / 128
    wA, wB = F.bcc_fields((N,) * 3, L, fA, xN)
    s = oracle.Solver("PmmmDCT", (N,) * 3)
    tr, _, _, phiA, phiB = s.scft(fA, ds, xN, L, 0.1, 2, F.extract(wA, "PmmmDCT"),
                                  F.extract(wB, "PmmmDCT"))
    assert abs(phiA.mean() - fA) < 1e-13
    assert abs(phiB.mean() - (1 - fA)) < 1e-13


def test_adjoint_identity_by_finite_differences(oracle):
    # d lnQ / dwA(r) = -(ds NsA / fA) phiA(r) / n  (Strang splitting, exact)
    N = 16
    fA, xN, L, ds = 0.28, 22.0, 4.0, 1 / 64  # NsA = 17 (truncated)
    wA, wB = F.bcc_fields((N,) * 3, L, fA, xN)
    wA, wB = F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT")
    s = oracle.Solver("PmmmDCT", (N,) * 3)
    _, _, _, phiA, _ = s.scft(fA, ds, xN, L, 0.0, 1, wA, wB)
    NsA = int(fA / ds)
    n = wA.size
    h = 2e-3
    for idx in [(0, 0, 0), (3, 5, 7), (7, 2, 1)]:
        def lnQ(d):
            w = wA.copy()
            w[idx] += d
            tr, *_ = s.scft(fA, ds, xN, L, 0.0, 1, w, wB)
            return -tr[0, 2]
        # fourth-order central difference
        fd = (-lnQ(2 * h) + 8 * lnQ(h) - 8 * lnQ(-h) + lnQ(-2 * h)) / (12 * h)
        an = -(ds * NsA / fA) * phiA[idx] / n
        assert abs(fd / an - 1) < 1e-6


def test_three_solvers_same_trace(oracle):
    # The reference's acceptance criterion (Readme.md:28-76), on seeded fields.
    c = CFG0
    wA, wB = F.bcc_fields((c["N"],) * 3, c["L"], c["fA"], c["xN"])
    traces = {}
    for name in ("PmmmDCT", "3m", "P1"):
        s = oracle.Solver(name, (c["N"],) * 3)
        tr, *_ = s.scft(c["fA"], c["ds"], c["xN"], c["L"], c["acc"], 3, F.extract(wA, name),
                        F.extract(wB, name))
        traces[name] = tr
    for name in ("3m", "P1"):
        assert rel_err(traces[name][:, 0], traces["PmmmDCT"][:, 0]) < 1e-12
        assert rel_err(traces[name][:, 5], traces["PmmmDCT"][:, 5]) < 1e-10


# ---- 3m with translational parts ----------------------------------------
# The reference's 3m solver with translational part g (solver_3m.h:172-255,
# generate_m3_operation solver_utilities.h:55-66) is the P1 diffusion of the
# symmetric full-grid field: its phase-folded coefficients are pinned by
# 3m_g(f) == P1(expand_g(f))[::2, ::2, ::2], with P1 pinned above by numpy
# fixtures.  Mirror-plane offsets (fractional parts on a mirrored axis) fix
# the direction of the translations: the other orientation fails them.
M3_G = [
    [0.5, 0.5, 0, 0, 0, 0, 0, 0, 0],           # n-glide normal to z
    [0, 0, 0, 0.5, 0, 0.5, 0, 0, 0],           # glide normal to y
    [0, 0, 0, 0, 0, 0, 0, 0.5, 0.5],           # glide normal to x
    [0.5, 0.5, 0, 0.5, 0, 0.5, 0, 0.5, 0.5],   # all three
    [0, 0, 0.25, 0, 0, 0, 0, 0, 0],            # mirror plane at z = 1/4 - h/2
    [0.5, 0, 0.25, 0, 0.125, 0.5, 0.25, 0, 0],
]


@pytest.mark.parametrize("g", M3_G)
def test_3m_translational_parts_equal_p1_on_expanded_fields(oracle, g):
    N, cell, ds = (16, 16, 32), (4.0, 4.3, 4.1), 0.02
    m3 = oracle.Solver("3m", N, g=g)
    m3.set_cell_para(cell)
    m3.set_contour_step(ds)
    p1 = oracle.Solver("P1", N)
    p1.set_cell_para(cell)
    p1.set_contour_step(ds)
    red = np.random.default_rng(17).standard_normal(m3.grid_size_phy)
    full = F.expand_3m(red, g)
    assert rel_err(m3.apply(red), p1.apply(full)[::2, ::2, ::2]) < 1e-13
    # the expanded result keeps the symmetry
    out = p1.apply(full)
    assert rel_err(F.expand_3m(out[::2, ::2, ::2], g), out) < 1e-13


def test_3m_expansion_rejects_invalid_operations():
    red = np.zeros((4, 4, 8))
    with pytest.raises(ValueError):
        F.expand_3m(red, [0.3, 0, 0, 0, 0, 0, 0, 0, 0])  # not whole grid points
    with pytest.raises(ValueError):
        F.expand_3m(red, [0, 0, 1 / 16, 0, 0, 0, 0, 0, 0])  # maps evens onto evens


def test_scft_golden_traces_agree_across_solvers(oracle):
    """tests/golden/scft_cfg1.npz (make_scft_traces.py): the three solvers'
    200-iteration traces on the 64^3 GPU workload agree to 1e-10 in F and
    MaxIncompressibility (the reference's acceptance criterion, Readme.md:28-76),
    and the first iteration is what the oracle computes now."""
    import os
    from conftest import GOLDEN

    g = np.load(os.path.join(GOLDEN, "scft_cfg1.npz"))
    p = g["PmmmDCT_trace"]
    assert p.shape == (200, 7)
    for name in ("3m", "P1"):
        t = g[f"{name}_trace"]
        assert np.abs(t[:, 0] / p[:, 0] - 1).max() < 1e-10
        assert np.abs(t[:, 5] / p[:, 5] - 1).max() < 1e-10
    N, fA, xN, L, ds = 64, 0.25, 20.0, 4.0, 0.005
    wA, wB = F.bcc_fields((N,) * 3, L, fA, xN)
    s = oracle.Solver("PmmmDCT", (N,) * 3)
    tr, *_ = s.scft(fA, ds, xN, L, 0.1, 1, F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT"))
    assert np.abs(tr[0] - p[0]).max() <= 1e-13 * np.abs(p[0]).max()


# ---- print_energy and calc_new_fields -------------------------------------
# print_energy (crysSCFT_sample.cpp:398-424) and calc_new_fields (:362-371)
# are pinned by what the mathematics fixes, not by their own formulas:
#   * uniform fields wA = a, wB = b: every trace column and the mixed fields
#     in closed form (q(s) q+(s) = Q at every s, so phiA = fA, phiB = fB);
#   * any fields: with H = <(wA-wB)^2/(4 xN) - (wA+wB)/2 + xN/4> - ln Q (no
#     density in it), t = 1-phiA-phiB, d = phiA-phiB and v = wA-wB, the four
#     printed terms satisfy the algebraic identity
#         F - H = <xN t^2 / 4> - <(v + xN d)^2 / (4 xN)>,
#     which a dropped or sign-flipped term, or phiA/phiB swapped in a term,
#     violates by O(the term);
#   * the mixing moves wA + wB by -acc xN t and wA - wB by -acc (v + xN d);
#   * at the fixed point of the mixing the printed F equals H built from Q.
def _uniform_case(oracle, name, fA, ds, a, b, xN, acc):
    s = oracle.Solver(name, (8, 8, 16))
    shape = s.grid_size_phy
    return s.scft(fA, ds, xN, 4.0, acc, 1, np.full(shape, a), np.full(shape, b))


@pytest.mark.parametrize("name", ["P1", "3m", "PmmmDCT"])
@pytest.mark.parametrize("fA,ds", [(0.3, 0.01), (0.28, 1 / 128)])  # NsA = 30, 35 (truncated)
def test_print_energy_uniform_closed_form(oracle, name, fA, ds):
    xN, a, b, acc = 17.0, 0.7, -0.4, 0.15
    tr, wA, wB, phiA, phiB = _uniform_case(oracle, name, fA, ds, a, b, xN, acc)
    fB = 1 - fA
    NsA, NsB = int(fA / ds), int(fB / ds)
    enth, mlnq, entr = xN * fA * fB, ds * (a * NsA + b * NsB), -(a * fA + b * fB)
    r = tr[0]
    assert abs(r[1] - enth) < 1e-13 * abs(enth)
    assert abs(r[2] - mlnq) < 1e-12 * abs(mlnq)
    assert abs(r[3] - entr) < 1e-12 * abs(entr)
    assert abs(r[4]) < 1e-13
    assert r[5] < 1e-13
    assert abs(r[0] - (enth + mlnq + entr)) < 1e-12 * abs(r[0])
    assert abs(r[6] - np.exp(-mlnq)) < 1e-13 * np.exp(-mlnq)
    # calc_new_fields with the pre-update ksi = (a + b - xN) / 2 (calc_ksi, :341-346)
    ksi = 0.5 * (a + b - xN)
    assert np.abs(wA - (a + acc * (xN * fB + ksi - a))).max() < 1e-13
    assert np.abs(wB - (b + acc * (xN * fA + ksi - b))).max() < 1e-13


def _noisy_fields(name, N, fA, xN, L, seed=3, amp=0.5):
    wA, wB = F.bcc_fields((N,) * 3, L, fA, xN)
    a, b = F.extract(wA, name), F.extract(wB, name)
    rng = np.random.default_rng(seed)
    return a + amp * rng.standard_normal(a.shape), b + amp * rng.standard_normal(b.shape)


def field_H(wA, wB, xN, Q):
    """The field-theoretic free energy: no density enters it."""
    return float(np.mean((wA - wB) ** 2 / (4 * xN) - (wA + wB) / 2 + xN / 4) - np.log(Q))


@pytest.mark.parametrize("name", ["P1", "3m", "PmmmDCT"])
def test_print_energy_identity_any_fields(oracle, name):
    N, fA, xN, L, ds = 16, 0.28, 22.0, 4.0, 1 / 64
    a, b = _noisy_fields(name, N, fA, xN, L)
    s = oracle.Solver(name, (N,) * 3)
    tr, wA, wB, phiA, phiB = s.scft(fA, ds, xN, L, 0.0, 1, a, b)
    r = tr[0]
    np.testing.assert_array_equal(wA, a)  # acceptance 0 leaves the fields
    t, d, v = 1 - phiA - phiB, phiA - phiB, a - b
    rhs = float(np.mean(xN * t ** 2 / 4) - np.mean((v + xN * d) ** 2 / (4 * xN)))
    assert abs((r[0] - field_H(a, b, xN, r[6])) - rhs) < 1e-13 * abs(r[0])
    # every term is far above that tolerance, so dropping or flipping any fails
    for k in (1, 2, 3, 4):
        assert abs(r[k]) > 1e-6
    assert r[0] == r[1] + r[2] + r[3] + r[4]
    assert r[2] == -np.log(r[6])
    assert abs(r[5] - np.abs(t).max()) < 1e-15
    assert r[5] > 1e-3


@pytest.mark.parametrize("name", ["P1", "PmmmDCT"])
def test_mixing_moves_the_fields_as_the_densities_fix(oracle, name):
    N, fA, xN, L, ds, acc = 16, 0.28, 22.0, 4.0, 1 / 64, 0.2
    a, b = _noisy_fields(name, N, fA, xN, L, seed=5)
    s = oracle.Solver(name, (N,) * 3)
    _, wA, wB, phiA, phiB = s.scft(fA, ds, xN, L, acc, 1, a, b)
    t, d, v = 1 - phiA - phiB, phiA - phiB, a - b
    scale = np.abs(a).max()
    assert np.abs((wA + wB) - (a + b) + acc * xN * t).max() < 1e-13 * scale
    assert np.abs((wA - wB) - v + acc * (v + xN * d)).max() < 1e-13 * scale


def test_converged_free_energy_equals_field_functional(oracle):
    """At the fixed point of simple mixing (wA = xN phiB + ksi, wB = xN phiA +
    ksi, hence phiA + phiB = 1) the printed F equals H[w], which is built from
    the fields and Q alone: an independent route to F."""
    N, fA, xN, L, ds, acc = 16, 0.25, 20.0, 4.0, 0.02, 0.5
    wA, wB = F.bcc_fields((N,) * 3, L, fA, xN)
    a, b = F.extract(wA, "PmmmDCT"), F.extract(wB, "PmmmDCT")
    s = oracle.Solver("PmmmDCT", (N,) * 3)
    tr, wA2, wB2, _, _ = s.scft(fA, ds, xN, L, acc, 600, a, b)
    assert tr[-1, 5] < 1e-10
    # the last iteration's fields: one more evaluation without mixing
    tr1, *_ = s.scft(fA, ds, xN, L, 0.0, 1, wA2, wB2)
    F_last, Q_last = tr1[0, 0], tr1[0, 6]
    assert abs(F_last - field_H(wA2, wB2, xN, Q_last)) < 1e-11 * abs(F_last)
    # and the structured phase is not the disordered melt
    assert abs(F_last - (xN * fA * (1 - fA))) > 1e-3
