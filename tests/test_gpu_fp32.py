"""The FP32 form of the step (arithmetic="fp32") against the FP64 form and the
oracle.

In the FP32 form the interface states, their characteristic reconstruction,
the Riemann fluxes and the cell E_z quadrature are FP32; the conserved state,
the face->cell collocation, the positivity recompute, the CFL rates and their
reduction, the corner E_z, the update and the conserved sums stay FP64.  It is
not a bit-parity form: its bar is a per-variable relative L1 distance from
the FP64 run (momentum and field components against their vector's norm).
The tolerance written here, 1e-4 for every variable, is the density bar the
FP32 form is held to at Orszag-Tang 4096^2 over 500 steps; these short runs
land orders of magnitude below it.  div B stays at round-off because the
field update is the discrete curl of an FP64 corner E_z."""
import numpy as np
import pytest

from gpu_helpers import interior, rel_l1_state
from oracle_ctypes import OracleSolver, ic_spec

pytestmark = pytest.mark.gpu

G53 = 5.0 / 3.0
NAMES = ("rho", "mx", "my", "mz", "energy", "bx_face", "by_face", "bz_cell")
FP32_L1_BAR = 1e-4


def make_solver(nx, ny, ng, gamma, planes, arithmetic, recon="mp7", fluid=("periodic",) * 4, lx=1.0, ly=1.0):
    from quasar_amd import mhd
    cfg = mhd.MhdConfig()
    cfg.grid = mhd.Grid2D(nx, ny, lx, ly, 0.0, 0.0, ng)
    cfg.gamma = gamma
    cfg.reconstruction = recon
    cfg.arithmetic = arithmetic
    for s in range(4):
        cfg.boundary.set_fluid_side(s, fluid[s])
        cfg.boundary.set_field_side(s, fluid[s])
    sol = mhd.MhdSolver2D(cfg)
    for k, name in enumerate(mhd.STATE_COMPONENTS):
        sol.seed_state(name, planes[k])
    return sol


def state(sol, nx, ny, ng):
    return [interior(sol.state_component_to_host(nm), nx, ny, ng) for nm in NAMES]


def ot_planes(oracle, n, ng, noise):
    g = oracle.grid(n, n, ng)
    spec = ic_spec("orszag_tang", G53, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0],
                   ot_rho=25 / (36 * np.pi), ot_p=5 / (12 * np.pi), noise_amplitude=noise, noise_seed=0)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    return g, planes


def test_arithmetic_option_is_validated(oracle):
    _, planes = ot_planes(oracle, 16, 4, 0.0)
    with pytest.raises(ValueError):
        make_solver(16, 16, 4, G53, planes, "fp16")


def test_fp32_orszag_tang_tracks_fp64(oracle):
    """Orszag-Tang 192^2 with the seeded 1e-3 perturbation, MP7, 30 auto-dt
    steps in each form."""
    n, ng, steps = 192, 4, 30
    _, planes = ot_planes(oracle, n, ng, 1e-3)
    f64 = make_solver(n, n, ng, G53, planes, "fp64")
    f32 = make_solver(n, n, ng, G53, planes, "fp32")
    d64, t64 = f64.run(steps)
    d32, t32 = f32.run(steps)
    assert d64 == steps and d32 == steps
    # dt comes from FP32 face states reduced in FP64.
    assert abs(t32 - t64) <= 1e-5 * t64
    a, b = state(f32, n, n, ng), state(f64, n, n, ng)
    errs = {nm: rel_l1_state(a, b, k) for k, nm in enumerate(NAMES)}
    print("fp32 vs fp64 rel L1:", errs)
    assert all(v <= FP32_L1_BAR for v in errs.values()), errs
    assert errs["rho"] > 0.0  # the forms really differ
    assert f32.divergence_b_max() <= 1e-11


def test_fp32_matches_oracle_on_a_small_grid(oracle):
    """Same bar against the oracle itself (fixed dt from the oracle)."""
    n, ng = 48, 4
    g, planes = ot_planes(oracle, n, ng, 1e-3)
    f32 = make_solver(n, n, ng, G53, planes, "fp32")
    ref = OracleSolver(oracle, g, G53, 0.4, 7)
    ref.seed(planes)
    for _ in range(6):
        dt = ref.cfl_limit()
        assert abs(f32.cfl_limit() - dt) <= 1e-5 * dt
        f32.step_unchecked(dt)
        ref.step_unchecked(dt)
    out = ref.state()
    exp = [interior(out[k], n, n, ng) for k in range(8)]
    got = state(f32, n, n, ng)
    errs = {nm: rel_l1_state(got, exp, k) for k, nm in enumerate(NAMES)}
    assert all(v <= FP32_L1_BAR for v in errs.values()), errs


def test_fp32_brio_wu_mp5_tracks_fp64(oracle):
    """Discontinuous data with outflow sides: the FP32 MP path, its fallback
    chain and the outflow ghost closures."""
    nx, ny, ng = 200, 4, 3
    g = oracle.grid(nx, ny, ng, 1.0, ny / nx)
    spec = ic_spec("brio_wu", 2.0, left=[1, 1, 0, 0, 0, 0.75, 1, 0], right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
    st, planes = oracle.seed(g, spec)
    assert st == 0
    sides = ("outflow", "outflow", "periodic", "periodic")
    f64 = make_solver(nx, ny, ng, 2.0, planes, "fp64", recon="mp5", fluid=sides, ly=ny / nx)
    f32 = make_solver(nx, ny, ng, 2.0, planes, "fp32", recon="mp5", fluid=sides, ly=ny / nx)
    d64, _ = f64.run(60)
    d32, _ = f32.run(60)
    assert d64 == d32 == 60
    a, b = state(f32, nx, ny, ng), state(f64, nx, ny, ng)
    errs = {nm: rel_l1_state(a, b, k) for k, nm in enumerate(NAMES)}
    print("brio-wu fp32 vs fp64 rel L1:", errs)
    assert all(v <= FP32_L1_BAR for v in errs.values()), errs
