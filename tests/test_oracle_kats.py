"""Known-answer tests and closed forms that pin the oracle independently of the
reference build: the reference's own KATs (tests/unit/numerics/
test_hlld_riemann.cpp, test_mp_limiter.cpp), exact-rational brute force for the
correctly rounded sums, Jacobian diagonalisation for the eigensystem, polynomial
exactness for the collocation and transverse rows, and the Random123 Philox
vectors."""
from fractions import Fraction

import numpy as np
import pytest

from oracle_ctypes import conserved


def analytic_flux_x(u, gamma):
    rho, mx, my, mz, e, bx, by, bz = u
    vx, vy, vz = mx / rho, my / rho, mz / rho
    p = (gamma - 1) * (e - 0.5 * rho * (vx * vx + vy * vy + vz * vz) - 0.5 * (bx * bx + by * by + bz * bz))
    pt = p + 0.5 * (bx * bx + by * by + bz * bz)
    vdb = vx * bx + vy * by + vz * bz
    return np.array([mx, mx * vx + pt - bx * bx, my * vx - bx * by, mz * vx - bx * bz,
                     (e + pt) * vx - bx * vdb, 0.0, by * vx - bx * vy, bz * vx - bx * vz])


# --- reference KATs (test_hlld_riemann.cpp) ----------------------------------

def test_hlld_rankine_hugoniot_known_answer(oracle):
    """test_hlld_riemann.cpp:357-377: right double-star energy flux."""
    g = 5.0 / 3.0
    L = conserved(0.5, 0.2, 0.8, -0.3, 0.7, 0.9, 0.4, 0.5, g)
    R = conserved(1.4, -0.5, -0.2, 0.9, 1.2, 0.9, -0.7, -0.1, g)
    out = oracle.hlld_flux_x(L, R, g)
    fR = analytic_flux_x(R, g)
    sR, saR = 2.1418875868531733, 0.3253072618499077
    eRs, eRss = 3.3054446710970824, 3.0613530843929193
    rh = fR[4] + sR * (eRs - R[4]) + saR * (eRss - eRs)
    assert abs(out[4] - rh) <= 2e-13
    assert abs(out[4] - (-2.0871013232600024)) <= 2e-13
    assert abs(out[4] - (-1.9282917918373412)) > 0.1  # sign-sensitive (Eq. 63)


def test_hlld_consistency_equals_analytic_flux(oracle):
    """test_hlld_riemann.cpp:132-152."""
    g = 5.0 / 3.0
    s = conserved(1.3, 0.4, -0.2, 0.3, 1.1, 0.6, 0.5, -0.4, g)
    np.testing.assert_allclose(oracle.hlld_flux_x(s, s, g), analytic_flux_x(s, g), atol=1e-12)


def test_hll_canonical_asymmetric_weights(oracle):
    """test_hlld_riemann.cpp:288-315."""
    g = 5.0 / 3.0
    sl, sr = -2.0, 5.0
    L = conserved(1.3, 0.7, -0.2, 0.1, 0.9, 0.4, 0.3, -0.1, g)
    R = conserved(0.6, -0.4, 0.5, -0.3, 1.2, 0.4, -0.2, 0.6, g)
    fl, fr = analytic_flux_x(L, g), analytic_flux_x(R, g)
    got = oracle.hll_flux_x(L, R, sl, sr, g)
    exp = (sr * fl - sl * fr + sl * sr * (R - L)) / (sr - sl)
    exp[5] = 0.0
    np.testing.assert_allclose(got, exp, atol=4e-15, rtol=0)


def test_hlld_unequal_normal_fields_normalised(oracle):
    """test_hlld_riemann.cpp:317-345."""
    g = 5.0 / 3.0
    L = conserved(1.0, 0.2, -0.1, 0.3, 0.9, 0.6, 0.4, -0.2, g)
    R = conserved(0.8, -0.15, 0.25, -0.1, 1.1, 1.0, -0.3, 0.5, g)
    bn = 0.5 * L[5] + 0.5 * R[5]
    Ln, Rn = L.copy(), R.copy()
    Ln[4] += 0.5 * bn * bn - 0.5 * Ln[5] ** 2
    Rn[4] += 0.5 * bn * bn - 0.5 * Rn[5] ** 2
    Ln[5] = Rn[5] = bn
    np.testing.assert_allclose(oracle.hlld_flux_x(L, R, g), oracle.hlld_flux_x(Ln, Rn, g),
                               atol=1e-12)


# --- reference KATs (test_mp_limiter.cpp) ------------------------------------

def test_mp_constant_and_steps(oracle):
    assert abs(oracle.mp5_reconstruct([3.5] * 5) - 3.5) <= 1e-12
    assert abs(oracle.mp7_reconstruct([3.5] * 7) - 3.5) <= 1e-12
    assert oracle.mp5_interp([0, 0, 1, 1, 1]) > 1 + 1e-6
    assert abs(oracle.mp5_reconstruct([0, 0, 1, 1, 1]) - 1.0) <= 1e-9
    assert oracle.mp5_interp([0, 0, 0, 0, 1]) < -1e-6
    assert abs(oracle.mp5_reconstruct([0, 0, 0, 0, 1])) <= 1e-9
    assert oracle.mp7_interp([0, 0, 0, 1, 1, 1, 1]) > 1 + 1e-6
    assert abs(oracle.mp7_reconstruct([0, 0, 0, 1, 1, 1, 1]) - 1.0) <= 1e-9


def test_mp_no_new_extremum_and_amplitude_invariance(oracle):
    v = np.array([1.0] * 6 + [0.0] * 6)
    for i in range(2, 10):
        lim = oracle.mp5_reconstruct(v[i - 2:i + 3])
        assert v[i - 2:i + 3].min() - 1e-12 <= lim <= v[i - 2:i + 3].max() + 1e-12
    a = 1e-100
    raw = oracle.mp5_interp([0, 0, a, a, a])
    lim = oracle.mp5_reconstruct([0, 0, a, a, a])
    assert raw > a and 0 <= lim <= a and abs(lim / a - 1) <= 1e-12


def test_mp_coefficients_are_finite_volume_exact(oracle):
    """The unlimited MP5/MP7 interpolants reproduce the face value of cell averages
    of polynomials of degree <= 4 / <= 6 (textbook FV reconstruction)."""
    rng = np.random.default_rng(0)
    for deg, n, f in ((4, 5, oracle.mp5_interp), (6, 7, oracle.mp7_interp)):
        c = rng.normal(size=deg + 1)
        P = np.polynomial.Polynomial(c).integ()
        h = n // 2
        avgs = [P(k + 0.5) - P(k - 0.5) for k in range(-h, h + 1)]  # cells centred at k
        assert abs(f(avgs) - np.polynomial.Polynomial(c)(0.5)) <= 1e-12 * max(1, abs(c).sum())


# --- exact sums --------------------------------------------------------------

def test_cr_sum_matches_exact_rational(oracle):
    rng = np.random.default_rng(1)
    for t in range(3000):
        n = int(rng.integers(1, 12))
        x = rng.normal(size=n) * 2.0 ** rng.integers(-60, 60, size=n)
        if n > 2 and t % 3 == 0:
            x[-1] = -float(sum(Fraction(v) for v in x[:-1]))  # near-total cancellation
        exact = sum(Fraction(float(v)) for v in x)
        assert oracle.cr_sum(x) == float(exact)


# --- eigensystem -------------------------------------------------------------

def _primitive_jacobian_conserved(u, gamma):
    """dF/dU of the 7 wave-carrying conserved components (normal B fixed),
    by central differences of the analytic flux."""
    idx = [0, 1, 2, 3, 4, 6, 7]
    J = np.zeros((7, 7))
    for a, ka in enumerate(idx):
        h = 1e-6 * max(1.0, abs(u[ka]))
        up, um = u.copy(), u.copy()
        up[ka] += h
        um[ka] -= h
        J[:, a] = (analytic_flux_x(up, gamma)[idx] - analytic_flux_x(um, gamma)[idx]) / (2 * h)
    return J


@pytest.mark.parametrize("seed", range(6))
def test_eigensystem_diagonalises_flux_jacobian(oracle, seed):
    """L R = I and R diag(lambda) L = dF/dU (test_mhd_eigensystem.cpp:201-509)."""
    rng = np.random.default_rng(seed)
    g = 5.0 / 3.0
    u = conserved(rng.uniform(0.5, 2), *rng.normal(size=3), rng.uniform(0.3, 2),
                  *rng.normal(size=3), g)
    L, R, s, fb = oracle.eigen_build(u, g)
    assert fb == 0
    np.testing.assert_allclose(L @ R, np.eye(7), atol=1e-12)
    np.testing.assert_allclose(R @ np.diag(s) @ L, _primitive_jacobian_conserved(u, g),
                               atol=2e-6, rtol=2e-6)


# --- collocation and transverse rows: polynomial exactness --------------------

@pytest.mark.parametrize("ng,deg", [(4, 7), (3, 5), (2, 3)])
def test_face_to_cell_collocation_is_polynomial_exact(oracle, ng, deg):
    """mhd_staggering.hpp:204-223: the 8/6/4-tap face->cell rule integrates
    polynomials of degree <= 7/5/3 exactly, including the one-sided closures."""
    nx, ny = 10, 6
    g = oracle.grid(nx, ny, ng, 1.0, 1.0, 0.0, 0.0)
    c = np.random.default_rng(ng).normal(size=deg + 1)
    P = np.polynomial.Polynomial(c)
    I = P.integ()
    face = np.zeros((ny + 2 * ng) * (nx + 2 * ng))
    for j in range(-ng, ny + ng):
        for i in range(-ng, nx + ng):
            face[(i + ng) + (nx + 2 * ng) * (j + ng)] = P(float(i))  # face i at x = i
    for i in range(-ng, nx + ng):
        exact = I(i + 1.0) - I(float(i))
        got = oracle.cell_bx(g, face, i, 0, 0)
        assert abs(got - exact) <= 1e-9 * max(1.0, abs(exact)), (i, got, exact)


@pytest.mark.parametrize("order,deg", [(7, 6), (5, 4)])
def test_transverse_point_rows_are_polynomial_exact(oracle, order, deg):
    """test_mhd_multidimensional_flux.cpp:169-230: point rows recover the pointwise
    value at each Gauss node from face averages of degree <= 6 / 4 polynomials."""
    nodes = {7: [-0.430568155797026286, -0.169990521792428156, 0.169990521792428156,
                 0.430568155797026286],
             5: [-0.387298334620741702, 0.0, 0.387298334620741702]}[order]
    weights = {7: [0.173927422568726925, 0.326072577431273103, 0.326072577431273103,
                   0.173927422568726925],
               5: [5 / 18, 4 / 9, 5 / 18]}[order]
    c = np.random.default_rng(order).normal(size=deg + 1)
    P = np.polynomial.Polynomial(c)
    I = P.integ()
    h = order // 2
    avgs = np.array([I(k + 0.5) - I(k - 0.5) for k in range(-h, h + 1)])
    for node, x in enumerate(nodes):
        assert abs(oracle.point_value(order, node, avgs) - P(x)) <= 1e-12 * np.abs(c).sum()
    # Gauss-Legendre average of a degree <= 2n-1 polynomial over the face is exact.
    assert abs(oracle.gauss_average(order, P(np.array(nodes))) * 1.0 - (I(0.5) - I(-0.5))) <= 1e-13
    assert abs(sum(weights) - 1.0) <= 1e-15


# --- Philox-4x32-10 ------------------------------------------------------------

PHILOX_KAT = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
              ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2,
               (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
              ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
               (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]


def test_philox_random123_vectors(oracle):
    """Random123 kat_vectors for philox4x32_10."""
    import ctypes as C
    for ctr, key, expect in PHILOX_KAT:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        oracle.lib.orc_philox4x32_10(c, k)
        assert tuple(c) == expect
    # The uniform mapping: 53 high bits of (word1:word0) onto [-1, 1).
    u = oracle.philox(0, 0, 0)
    bits = PHILOX_KAT[0][2][0] | (PHILOX_KAT[0][2][1] << 32)
    assert u == 2.0 * ((bits >> 11) * 2.0 ** -53) - 1.0
