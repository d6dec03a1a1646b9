"""Pin the oracle's per-state numerics bit for bit against the reference's own
header-only numerics, compiled in place into oracle/_ref (see oracle/Makefile).

Covers: exact product sums (scaled_arithmetic.hpp), EOS (mhd_state.hpp),
HLLD/HLL/LF and the physical flux (hlld_core.hpp), MP5/MP7 (mp_limiter.hpp),
the 7-wave eigensystem and characteristic projection (mhd_eigensystem.hpp,
characteristic_projection.hpp), face->cell collocation including the one-sided
ghost closures (mhd_staggering.hpp) and the transverse Gauss rows
(finite_volume_quadrature.hpp)."""
import numpy as np
import pytest

from oracle_ctypes import DP, conserved

GAMMAS = (5.0 / 3.0, 1.4, 2.0)


def bits_equal(a, b):
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    return np.array_equal(a.view(np.int64), b.view(np.int64))


def random_state(rng, gamma, extreme=False):
    rho = rng.uniform(0.05, 3.0) * (10.0 ** rng.integers(-3, 3) if extreme else 1.0)
    v = rng.normal(0, 1.5, 3)
    p = rng.uniform(0.01, 3.0) * (10.0 ** rng.integers(-4, 2) if extreme else 1.0)
    b = rng.normal(0, 1.5, 3)
    if rng.uniform() < 0.1:
        b[0] = 0.0
    if rng.uniform() < 0.1:
        b[1:] = 0.0
    return conserved(rho, *v, p, *b, gamma)


@pytest.mark.parametrize("gamma", GAMMAS)
def test_riemann_solvers_bitwise(oracle, reference, gamma):
    rng = np.random.default_rng(11)
    for t in range(3000):
        L = random_state(rng, gamma, extreme=t % 5 == 0)
        R = random_state(rng, gamma, extreme=t % 7 == 0)
        if t % 3 == 0:
            R[5] = L[5]
        if t % 11 == 0:
            R = L.copy()
        assert bits_equal(oracle.hlld_flux_x(L, R, gamma),
                          reference.call8("ref_hlld_flux_x", L, R, gamma)), t
        assert bits_equal(oracle.lf_flux_x(L, R, gamma),
                          reference.call8("ref_lf_flux_x", L, R, gamma)), t
        sl, sr = -abs(rng.normal()) * 3, abs(rng.normal()) * 3
        assert bits_equal(oracle.hll_flux_x(L, R, sl, sr, gamma),
                          reference.call8("ref_hll_flux_x", L, R, sl, sr, gamma)), t
        assert bits_equal(oracle.physical_flux_x(L, gamma),
                          reference.call8("ref_physical_flux_x", L, gamma)), t
        assert bits_equal(oracle.pressure(L, gamma), reference.scalar("ref_pressure", L, gamma))
        for d in (0, 1):
            assert bits_equal(oracle.fast_speed(L, d, gamma),
                              reference.scalar("ref_fast_speed", L, d, gamma))


def test_hlld_paths_all_exercised(oracle):
    """The random sweep above must reach every HLLD branch it claims to pin."""
    rng = np.random.default_rng(5)
    seen = set()
    g = 5.0 / 3.0
    for t in range(4000):
        L = random_state(rng, g, extreme=t % 5 == 0)
        R = random_state(rng, g)
        if t % 3 == 0:
            R[5] = L[5]
        _, path = oracle.hlld_flux_x(L, R, g, return_path=True)
        seen.add(path)
    assert {0, 2, 3} <= seen


def test_product_sums_bitwise(oracle, reference):
    rng = np.random.default_rng(3)
    for t in range(5000):
        n = int(rng.integers(1, 9))
        a = rng.normal(size=n) * 10.0 ** rng.integers(-8, 8, size=n)
        b = rng.normal(size=n) * 10.0 ** rng.integers(-8, 8, size=n)
        if t % 4 == 0 and n >= 2:  # planted near-cancellation
            a[1] = -a[0] * b[0] / b[1] * (1 + 1e-15 * rng.normal())
        assert bits_equal(oracle.product_sum(a, b),
                          reference.scalar("ref_product_sum", a, b, n)), (a, b)


def test_mp_limiter_bitwise(oracle, reference):
    rng = np.random.default_rng(4)
    for t in range(20000):
        v = rng.normal(size=7)
        kind = t % 4
        if kind == 1:
            v = np.cumsum(np.abs(v))
        elif kind == 2:
            v = np.where(rng.uniform(size=7) < 0.5, 0.0, 1.0) * rng.uniform(0.5, 2)
        elif kind == 3:
            v = np.sin(np.linspace(0, rng.uniform(0.1, 3), 7) + rng.uniform(0, 6))
        assert bits_equal(oracle.mp7_reconstruct(v), reference.scalar("ref_mp7_reconstruct", v))
        assert bits_equal(oracle.mp5_reconstruct(v[:5]),
                          reference.scalar("ref_mp5_reconstruct", v[:5]))
        assert bits_equal(oracle.mp7_interp(v), reference.scalar("ref_mp7_interp", v))
        assert bits_equal(oracle.mp5_interp(v[:5]), reference.scalar("ref_mp5_interp", v[:5]))
        vl = rng.normal()
        assert bits_equal(oracle.mp_limit(vl, v[:5]), reference.scalar("ref_mp_limit", vl, v[:5]))


@pytest.mark.parametrize("gamma", GAMMAS)
def test_eigensystem_and_projection_bitwise(oracle, reference, gamma):
    rng = np.random.default_rng(6)
    for t in range(2000):
        u = random_state(rng, gamma, extreme=t % 4 == 0)
        Lo, Ro, so, _ = oracle.eigen_build(u, gamma)
        Lr, Rr, sr = reference.eigen_build(u, gamma)
        assert bits_equal(Lo, Lr) and bits_equal(Ro, Rr) and bits_equal(so, sr), t
        d = rng.normal(size=8)
        w_ref = np.zeros(7)
        reference.lib.ref_to_char(d.ctypes.data_as(DP), u.ctypes.data_as(DP), gamma,
                                  w_ref.ctypes.data_as(DP))
        w_orc = np.zeros(7)
        oracle.lib.orc_to_char(d.ctypes.data_as(DP), Lo.ravel().copy().ctypes.data_as(DP),
                               w_orc.ctypes.data_as(DP))
        assert bits_equal(w_orc, w_ref)


@pytest.mark.parametrize("ng,nx,ny", [(4, 12, 9), (3, 10, 6), (2, 7, 5)])
def test_collocation_bitwise_including_one_sided_closures(oracle, reference, ng, nx, ny):
    rng = np.random.default_rng(7)
    g = oracle.grid(nx, ny, ng, 1.3, 0.7, 0.1, -0.2)
    n = (nx + 2 * ng) * (ny + 2 * ng)
    for trial in range(3):
        face = rng.normal(size=n) if trial else np.full(n, 1.25)
        fp = face.ctypes.data_as(DP)
        for j in range(-ng, ny + ng):
            for i in range(-ng, nx + ng):
                for co in (0, 1):
                    assert bits_equal(
                        oracle.cell_bx(g, face, i, j, co),
                        reference.lib.ref_cell_bx(nx, ny, ng, 1.3, 0.7, 0.1, -0.2, fp, i, j, co))
                    assert bits_equal(
                        oracle.cell_by(g, face, i, j, co),
                        reference.lib.ref_cell_by(nx, ny, ng, 1.3, 0.7, 0.1, -0.2, fp, i, j, co))


def test_transverse_rows_bitwise(oracle, reference):
    rng = np.random.default_rng(8)
    for t in range(3000):
        s7 = rng.normal(size=7) if t % 9 else np.full(7, 0.3)
        for node in range(4):
            assert bits_equal(oracle.point_value(7, node, s7),
                              reference.lib.ref_point_value7(node, s7.ctypes.data_as(DP)))
        for node in range(3):
            assert bits_equal(oracle.point_value(5, node, s7[:5]),
                              reference.lib.ref_point_value5(node, s7[:5].copy().ctypes.data_as(DP)))
        for order, count in ((5, 3), (7, 4)):
            q = s7[:count].copy()
            assert bits_equal(oracle.gauss_average(order, q),
                              reference.lib.ref_gauss_average(order, q.ctypes.data_as(DP)))
