"""Invariants that pin the oracle's composite kernels (reconstruction, face
quadrature, CT, SSP-RK3, CFL, positivity, ghosts, seeding) — the reference's own
solver-level checks (tests/unit/physics/mhd/*, tests/unit/numerics/test_ct_divergence.cpp,
test_ssprk3_integrator.cpp, test_mhd_device_compute.cpp, test_mhd_boundary.cpp)
restated as properties, since no golden trajectories exist in the reference."""
import numpy as np
import pytest

from oracle_ctypes import OracleSolver, conserved, ic_spec

G53 = 5.0 / 3.0


def padded(nx, ny, ng):
    return (nx + 2 * ng) * (ny + 2 * ng)


def idx(nx, ng, i, j):
    return (i + ng) + (nx + 2 * ng) * (j + ng)


def interior(a, nx, ny, ng):
    return a.reshape(ny + 2 * ng, nx + 2 * ng)[ng:ng + ny, ng:ng + nx]


def ot_state(oracle, n, ng=4, noise=0.0):
    g = oracle.grid(n, n, ng)
    spec = ic_spec("orszag_tang", G53, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0],
                   ot_rho=25 / (36 * np.pi), ot_p=5 / (12 * np.pi), noise_amplitude=noise,
                   noise_seed=0)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    return g, planes


@pytest.mark.parametrize("order,ng", [(7, 4), (5, 3), (2, 2)])
def test_free_stream_residual_is_exactly_zero(oracle, order, ng):
    nx, ny = 9, 7
    g = oracle.grid(nx, ny, ng)
    u = conserved(1.3, 0.4, -0.7, 0.2, 0.9, 0.6, -0.35, 0.25, G53)
    planes = [np.full(padded(nx, ny, ng), v) for v in u]
    s = OracleSolver(oracle, g, G53, 0.4, order)
    s.seed(planes)
    dudt = s.compute_residual()
    for c in range(8):
        assert np.all(interior(dudt[c], nx, ny, ng) == 0.0), c


def test_ssprk3_fixed_point_of_constant_state(oracle):
    """test_ssprk3_integrator.cpp:120-137: unchanged to round-off (the Shu-Osher
    weights 3/4, 1/3, 2/3 are rounded, so a few ulp are allowed)."""
    nx, ny, ng = 8, 8, 4
    g = oracle.grid(nx, ny, ng)
    u = conserved(0.8, -0.3, 0.2, 0.0, 1.2, 0.5, 0.1, 0.0, G53)
    planes = [np.full(padded(nx, ny, ng), v) for v in u]
    s = OracleSolver(oracle, g, G53, 0.4, 7)
    s.seed(planes)
    for _ in range(2):
        s.step_unchecked(s.cfl_limit())
    out = s.state()
    for c in range(8):
        assert np.max(np.abs(interior(out[c], nx, ny, ng) - u[c])) <= 4e-16 * max(1.0, abs(u[c]))


def test_orszag_tang_conserves_totals_and_div_b(oracle):
    n = 24
    g, planes = ot_state(oracle, n)
    s = OracleSolver(oracle, g, G53, 0.4, 7)
    s.seed(planes)
    m0, e0 = s.conserved_totals()
    mom0 = [interior(planes[c], n, n, 4).sum() for c in (1, 2)]
    for _ in range(2):
        s.step_unchecked(s.cfl_limit())
    m, e = s.conserved_totals()
    out = s.state()
    assert abs(m - m0) <= 1e-14 * abs(m0)
    assert abs(e - e0) <= 1e-14 * abs(e0)
    for c, ref in zip((1, 2), mom0):
        assert abs(interior(out[c], n, n, 4).sum() - ref) <= 1e-13
    # CT keeps div B at round-off (test_mhd_divergence_free.cpp:113-135).
    assert s.divergence_b_max() < 1e-10


def test_one_dimensional_data_stays_one_dimensional(oracle):
    """Brio-Wu on a periodic-y strip stays one-dimensional.  Not bitwise: the
    reference's one-sided ghost-row collocation (mhd_staggering.hpp:111-173) does
    not reproduce a constant exactly (checked bit for bit against the reference in
    test_oracle_vs_reference.py), so rows adjacent to the transverse halo differ
    from the interior rows by a few ulp."""
    nx, ny, ng = 40, 4, 3
    g = oracle.grid(nx, ny, ng, 1.0, ny / nx)
    spec = ic_spec("brio_wu", 2.0, left=[1, 1, 0, 0, 0, 0.75, 1, 0],
                   right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
    st, planes = oracle.seed(g, spec)
    assert st == 0
    s = OracleSolver(oracle, g, 2.0, 0.4, 5, fluid=(1, 1, 0, 0))
    s.seed(planes)
    for _ in range(5):
        s.step_unchecked(s.cfl_limit())
    out = s.state()
    for c in range(8):
        rows = interior(out[c], nx, ny, ng)
        assert np.max(np.abs(rows - rows[0:1, :])) <= 1e-13, c
    assert s.divergence_b_max() <= 1e-12


def test_transpose_covariance(oracle):
    """Transposing x<->y (with mx<->my, bx<->by) commutes with a step up to the
    rounding of the order in which tensor quadratures are evaluated."""
    n, ng = 16, 4
    g, planes = ot_state(oracle, n)
    s1 = OracleSolver(oracle, g, G53, 0.4, 7)
    s1.seed(planes)

    def T(a):
        return np.ascontiguousarray(a.reshape(n + 2 * ng, n + 2 * ng).T).ravel()

    tp = [T(planes[0]), T(planes[2]), T(planes[1]), T(planes[3]), T(planes[4]),
          T(planes[6]), T(planes[5]), T(planes[7])]
    s2 = OracleSolver(oracle, g, G53, 0.4, 7)
    s2.seed(tp)
    dt = min(s1.cfl_limit(), s2.cfl_limit())
    assert abs(s1.cfl_limit() - s2.cfl_limit()) <= 1e-15 * dt
    s1.step_unchecked(dt)
    s2.step_unchecked(dt)
    a, b = s1.state(), s2.state()
    back = [T(b[0]), T(b[2]), T(b[1]), T(b[3]), T(b[4]), T(b[6]), T(b[5]), T(b[7])]
    for c in range(8):
        x = interior(a[c], n, n, ng)
        y = interior(back[c], n, n, ng)
        assert np.max(np.abs(x - y)) <= 1e-13 * max(1.0, np.max(np.abs(x))), c


def test_cfl_hand_computed_uniform_state(oracle):
    """test_mhd_device_compute.cpp:512-900: uniform state, additive x+y rate,
    MP Courant factor clamped to 0.2."""
    nx, ny, ng = 6, 5, 4
    lx, ly = 1.0, 0.5
    g = oracle.grid(nx, ny, ng, lx, ly)
    rho, vx, vy, p, bx, by, bz = 1.1, 0.3, -0.45, 0.8, 0.6, 0.2, -0.1
    u = conserved(rho, vx, vy, 0.0, p, bx, by, bz, G53)
    planes = [np.full(padded(nx, ny, ng), v) for v in u]
    s = OracleSolver(oracle, g, G53, 0.4, 7)
    s.seed(planes)

    def cf(bn):
        a2 = G53 * p / rho
        b2 = (bx * bx + by * by + bz * bz) / rho
        return np.sqrt(0.5 * (a2 + b2 + np.sqrt((a2 + b2) ** 2 - 4 * a2 * bn * bn / rho)))

    rate = (abs(vx) + cf(bx)) / (lx / nx) + (abs(vy) + cf(by)) / (ly / ny)
    assert abs(s.cfl_limit() - 0.2 / rate) <= 1e-14 * (0.2 / rate)


def test_admissible_fraction_closed_form(oracle):
    """Density-limited convex fraction: theta = rho_a / (rho_a - rho_b)."""
    nx, ny, ng = 4, 3, 4
    g = oracle.grid(nx, ny, ng)
    a = conserved(1.0, 0, 0, 0, 1.0, 0.1, 0.0, 0.0, G53)
    b = a.copy()
    b[0] = -0.5
    base = [np.full(padded(nx, ny, ng), v) for v in a]
    cand = [np.full(padded(nx, ny, ng), v) for v in b]
    theta = oracle.admissible_fraction(g, base, cand)
    assert 0.0 < theta < 1.0
    assert abs(theta - 1.0 / 1.5) <= 1e-15 or theta <= 1.0 / 1.5
    same = oracle.admissible_fraction(g, base, base)
    assert same == 1.0


@pytest.mark.parametrize("modes", [(0, 0, 0, 0), (1, 1, 1, 1), (2, 2, 2, 2), (1, 1, 0, 0)])
def test_ghost_fill_index_maps(oracle, modes):
    """mhd_boundary_hip.hip rules (a)-(c), restated independently in numpy."""
    nx, ny, ng = 7, 6, 3
    g = oracle.grid(nx, ny, ng)
    rng = np.random.default_rng(3)
    u = [rng.normal(size=padded(nx, ny, ng)) for _ in range(8)]
    ref = [a.reshape(ny + 2 * ng, nx + 2 * ng).copy() for a in u]
    oracle.fill_ghosts(g, modes, modes, u)
    J = lambda j: j + ng  # noqa: E731
    I = J
    for side in range(4):
        mode = modes[side]
        x_side = side < 2
        lo = side in (0, 2)
        n = nx if x_side else ny
        lines = range(-ng, (ny if x_side else nx) + ng)
        for layer in range(1, ng + 1):
            gi = -layer if lo else n - 1 + layer
            si = (n - layer if lo else layer - 1) if mode == 0 else \
                 (0 if lo else n - 1) if mode == 1 else (layer - 1 if lo else n - layer)
            for line in lines:
                if x_side:
                    d, s_ = (J(line), I(gi)), (J(line), I(si))
                else:
                    d, s_ = (J(gi), I(line)), (J(si), I(line))
                sn = -1.0 if mode == 2 else 1.0
                for c in (0, 3, 4):
                    ref[c][d] = ref[c][s_]
                ref[1][d] = (sn if x_side else 1.0) * ref[1][s_]
                ref[2][d] = (1.0 if x_side else sn) * ref[2][s_]
                tang, norm = (6, 5) if x_side else (5, 6)
                ref[tang][d] = ref[tang][s_]
                ref[7][d] = ref[7][s_]
                nf = ref[norm]
                if mode == 2:
                    if x_side:
                        if lo:
                            nf[J(line), I(0)] = 0.0
                            nf[J(line), I(-layer)] = -nf[J(line), I(layer)]
                        elif layer == 1:
                            nf[J(line), I(nx)] = 0.0
                        else:
                            nf[J(line), I(nx + layer - 1)] = -nf[J(line), I(nx - layer + 1)]
                    else:
                        if lo:
                            nf[J(0), I(line)] = 0.0
                            nf[J(-layer), I(line)] = -nf[J(layer), I(line)]
                        elif layer == 1:
                            nf[J(ny), I(line)] = 0.0
                        else:
                            nf[J(ny + layer - 1), I(line)] = -nf[J(ny - layer + 1), I(line)]
                elif mode == 1 and not lo:
                    if layer > 1:
                        if x_side:
                            nf[J(line), I(nx - 1 + layer)] = nf[J(line), I(nx)]
                        else:
                            nf[J(ny - 1 + layer), I(line)] = nf[J(ny), I(line)]
                else:
                    nf[d] = nf[s_]
    for c in range(8):
        np.testing.assert_array_equal(u[c].reshape(ny + 2 * ng, nx + 2 * ng), ref[c])


def test_orszag_tang_seed_matches_closed_form(oracle):
    n, ng = 16, 4
    g, planes = ot_state(oracle, n)
    rho0, p0, b0 = 25 / (36 * np.pi), 5 / (12 * np.pi), 1 / np.sqrt(4 * np.pi)
    x = (np.arange(-ng, n + ng) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    np.testing.assert_allclose(planes[0].reshape(X.shape), rho0, rtol=0, atol=0)
    np.testing.assert_allclose(planes[1].reshape(X.shape), -rho0 * np.sin(2 * np.pi * Y),
                               atol=4e-16)
    np.testing.assert_allclose(planes[6].reshape(X.shape), b0 * np.sin(4 * np.pi * X), atol=4e-16)
    # Energy uses the collocated (8-tap) magnetic energy, not the raw face samples.
    bxc = np.array([oracle.cell_bx(g, planes[5], i, j) for j in range(n) for i in range(n)])
    byc = np.array([oracle.cell_by(g, planes[6], i, j) for j in range(n) for i in range(n)])
    e = interior(planes[4], n, n, ng).ravel()
    kin = 0.5 * (interior(planes[1], n, n, ng) ** 2 + interior(planes[2], n, n, ng) ** 2).ravel() / rho0
    np.testing.assert_allclose(e, p0 / (G53 - 1) + kin + 0.5 * (bxc ** 2 + byc ** 2), rtol=1e-14)


def test_confined_blob_seed_matches_closed_form(oracle):
    """Square blob (mhd_initial_conditions.hip:190-207): rho_in/p_in where both
    |x - cx| and |y - cy| are within blob_half (edges included), rho_out/p_out
    elsewhere, at rest, with zero in-plane field and a uniform bz."""
    n, ng, gamma, bz = 16, 4, 1.4, 0.3
    g = oracle.grid(n, n, ng)
    spec = ic_spec("confined_blob", gamma, center=[0.5, 0.5], blob_half=0.25, b_uniform=[0, 0, bz],
                   rho_in=10.0, rho_out=1.0, p_in=2.0, p_out=0.1)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    x = (np.arange(-ng, n + ng) + 0.5) / n
    X, Y = np.meshgrid(x, x)
    inside = (np.abs(X - 0.5) <= 0.25) & (np.abs(Y - 0.5) <= 0.25)
    assert 0 < inside.sum() < inside.size
    np.testing.assert_array_equal(planes[0].reshape(X.shape), np.where(inside, 10.0, 1.0))
    for c in (1, 2, 3, 5, 6):
        assert not planes[c].any()
    np.testing.assert_array_equal(planes[7], bz)
    p = np.where(inside, 2.0, 0.1)
    np.testing.assert_allclose(planes[4].reshape(X.shape), p / (gamma - 1) + 0.5 * bz * bz, rtol=1e-15)


def test_noise_perturbation_is_seeded_and_bounded(oracle):
    n = 12
    _, clean = ot_state(oracle, n)
    _, noisy = ot_state(oracle, n, noise=1e-3)
    _, again = ot_state(oracle, n, noise=1e-3)
    for c in range(8):
        np.testing.assert_array_equal(noisy[c], again[c])
    rel = interior(noisy[0], n, n, 4) / interior(clean[0], n, n, 4) - 1.0
    assert np.all(np.abs(rel) <= 1e-3) and np.std(rel) > 1e-4


def test_invalid_seeds_are_rejected(oracle):
    n, ng = 8, 4
    g, planes = ot_state(oracle, n)
    s = OracleSolver(oracle, g, G53, 0.4, 7)
    bad = [p.copy() for p in planes]
    bad[4][idx(n, ng, 3, 3)] = 0.0  # negative pressure
    s.seed(bad)
    with pytest.raises(ValueError):
        s.cfl_limit()
    s2 = OracleSolver(oracle, g, G53, 0.4, 7)
    nonsol = [p.copy() for p in planes]
    nonsol[5][idx(n, ng, 3, 3)] += 0.1  # non-solenoidal face field
    s2.seed(nonsol)
    with pytest.raises(ValueError):
        s2.cfl_limit()


def test_positivity_retry_keeps_state_admissible_and_conservative(oracle):
    """A strong blast on a coarse MP7 grid: whatever the controller does, the
    accepted state stays positive and the totals are conserved
    (test_mhd_positivity_preservation.cpp:353-520)."""
    n, ng = 16, 4
    g = oracle.grid(n, n, ng)
    spec = ic_spec("blast", 1.4, rho_ambient=1.0, p_ambient=0.1, p_core=1000.0,
                   r_in=0.12, center=[0.5, 0.5], b_uniform=[0.5, 0.0, 0.0])
    st, planes = oracle.seed(g, spec)
    assert st == 0
    s = OracleSolver(oracle, g, 1.4, 0.4, 7)
    s.seed(planes)
    m0, e0 = s.conserved_totals()
    for _ in range(3):
        s.step_unchecked(s.cfl_limit())
    m, e = s.conserved_totals()
    assert abs(m - m0) <= 1e-13 * m0 and abs(e - e0) <= 1e-13 * e0
    out = s.state()
    assert np.all(interior(out[0], n, n, ng) > 0)


def test_brio_wu_reference_fixture_invariants():
    """The committed oracle run of the Brio-Wu benchmark configuration
    (tests/golden/make_brio_wu_reference.py) against what the physics fixes:
    by t = 0.1 no wave has left [0,1], so mass and energy keep their initial
    totals; a one-dimensional run keeps bx = 0.75 on every face, no z
    components and equal rows (to the round-off the one-sided ghost-row
    collocation leaves); the far states keep the initial plateaus."""
    from pathlib import Path
    path = Path(__file__).resolve().parent / "golden" / "brio_wu_800x4_mp5_t0.1.npz"
    ref = np.load(path, allow_pickle=False)
    rho, mx, my, mz, e, bx, by, bz = ref["interior"]
    assert ref["interior"].shape == (8, 4, 800)
    assert float(ref["final_time"][0]) == 0.1 and len(ref["dts"]) == int(ref["steps"][0])
    assert abs(float(np.sum(ref["dts"])) - 0.1) <= 1e-15
    # Totals to the round-off of ~3000 steps of flux differences.
    assert abs(rho.mean() - 0.5625) <= 1e-12
    e0 = 0.5 * (1.0 + 0.5 * (0.75 ** 2 + 1.0)) + 0.5 * (0.1 + 0.5 * (0.75 ** 2 + 1.0))
    assert abs(e.mean() - e0) <= 1e-12
    assert np.max(np.abs(bx - 0.75)) <= 1e-12
    assert np.max(np.abs(mz)) <= 1e-14 and np.max(np.abs(bz)) <= 1e-14
    for a in ref["interior"]:
        assert np.max(np.abs(a - a[0])) <= 1e-12 * max(1.0, np.max(np.abs(a)))
    assert np.allclose(rho[0, :20], 1.0, atol=1e-12) and np.allclose(rho[0, -20:], 0.125, atol=1e-12)
    assert float(ref["divb"][0]) <= 1e-10
