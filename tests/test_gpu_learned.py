"""The learned form on the device, through the C ABI: the BF16-MFMA network
(forward and gradient) against the oracle's network with the same operand
rounding, training with the library's kernels, features and labels against
the oracle, the indicator kernel against the network, and the learned solver
against the plain one.

Tolerances: features are computed in the same order of operations as the
oracle and agree bit for bit.  The network accumulates its BF16 products in
FP32 (the oracle in FP64), and a hidden activation that lands on the other
side of a BF16 rounding boundary moves a logit by up to ~1e-3, so logits are
compared by median (<= 1e-5) and by the share of samples off by more than 2e-3
(<= 1%); gradients to 2e-3 relative L2 per parameter block.  Labels come from
branch decisions (MP limiter clamps, fallbacks, a flux-gap threshold) that
round-off can flip at isolated faces: >= 99.5% of cells agree."""
import ctypes as C
import pathlib

import numpy as np
import pytest

from gpu_helpers import ptrs, qgrid, scratch, to_dev, to_host, torch_cuda
from oracle_ctypes import ic_spec

pytestmark = pytest.mark.gpu

P = 4321
BLOCKS = {"W1": (0, 3200), "b1": (3200, 3232), "W2": (3232, 4256), "b2": (4256, 4288), "w3": (4288, 4320),
          "b3": (4320, 4321)}
ROOT = pathlib.Path(__file__).resolve().parents[1]


def lib():
    from quasar_amd import _lib
    return _lib.load()


def check(status):
    from quasar_amd import _lib
    _lib.check(status)


def rng(seed):
    return np.random.Generator(np.random.Philox(key=seed))


def vp(t):
    return C.c_void_p(t.data_ptr())


def packed_of(params_dev):
    torch = torch_cuda()
    packed = torch.zeros(5632, dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    check(lib().qmhd_mlp_pack(vp(params_dev), vp(packed), None))
    return packed


def gpu_logits(params, x):
    torch = torch_cuda()
    p = torch.from_numpy(params).cuda()
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    out = torch.zeros(x.shape[0], dtype=torch.float32, device="cuda")
    packed = packed_of(p)
    check(lib().qmhd_mlp_forward(vp(packed), vp(p), vp(xd), x.shape[0], vp(out), None))
    check(lib().qmhd_stream_synchronize(None))
    return out.cpu().numpy().astype(np.float64)


def gpu_grad(params, x, y, w, chunks=2):
    torch = torch_cuda()
    p = torch.from_numpy(params).cuda()
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    yd = torch.from_numpy(np.ascontiguousarray(y, np.float32)).cuda()
    wd = torch.from_numpy(np.ascontiguousarray(w, np.float32)).cuda()
    n = x.shape[0]
    part = torch.zeros(lib().qmhd_mlp_partial_floats(n, chunks), dtype=torch.float32, device="cuda")
    grad = torch.zeros(P + 1, dtype=torch.float32, device="cuda")
    packed = packed_of(p)
    check(lib().qmhd_mlp_gradient(vp(packed), vp(p), vp(xd), vp(yd), vp(wd), None, n, chunks, vp(part), vp(grad),
                                  None))
    check(lib().qmhd_stream_synchronize(None))
    g = grad.cpu().numpy().astype(np.float64)
    return g[P], g[:P]


def random_problem(seed, n):
    r = rng(seed)
    params = (r.normal(0, 0.25, P)).astype(np.float32)
    x = r.uniform(-1, 1, (n, 100)).astype(np.float32)
    y = (r.uniform(0, 1, n) < 0.3).astype(np.float32)
    w = r.uniform(0.5, 1.5, n).astype(np.float32) / n
    return params, x, y, w


def test_mlp_forward_matches_oracle(oracle):
    params, x, _, _ = random_problem(11, 2003)  # not a multiple of 16
    got = gpu_logits(params, x)
    exp = oracle.mlp_forward(params, x, bf16=True)
    err = np.abs(got - exp)
    assert np.median(err) <= 1e-5 * (1 + np.median(np.abs(exp)))
    assert (err > 2e-3 * (1 + np.abs(exp))).mean() <= 0.01, err.max()


def test_mlp_gradient_matches_oracle(oracle):
    params, x, y, w = random_problem(12, 1000)
    loss_g, g = gpu_grad(params, x, y, w)
    loss_o, go = oracle.mlp_loss_grad(params, x, y, w, bf16=True)
    assert abs(loss_g - loss_o) <= 1e-4 * abs(loss_o)
    for name, (a, b) in BLOCKS.items():
        num = np.linalg.norm(g[a:b] - go[a:b])
        den = np.linalg.norm(go[a:b]) + 1e-30
        assert num <= 2e-3 * den, (name, num / den)


def test_mlp_gradient_is_deterministic():
    params, x, y, w = random_problem(13, 700)
    l1, g1 = gpu_grad(params, x, y, w, chunks=3)
    l2, g2 = gpu_grad(params, x, y, w, chunks=3)
    assert l1 == l2 and np.array_equal(g1, g2)


def test_training_with_device_kernels_learns_a_rule():
    """Adam steps of the library's kernels on a separable synthetic rule."""
    from quasar_amd import learned
    r = rng(14)
    x = r.uniform(-1, 1, (8192, 100)).astype(np.float32)
    y = ((x[:, 12] + 0.5 * x[:, 60] - 0.3 * x[:, 99]) > 0.1).astype(np.float32)
    tr = learned.Trainer(x, y, learned.init_params(3), pos_weight=1.0, seed=5)
    first = tr.epoch(512, 3e-3, 0)
    for e in range(1, 12):
        last = tr.epoch(512, 3e-3, e)
    z = learned.device_logits(tr.params(), x)
    acc = ((z >= 0) == (y > 0.5)).mean()
    assert last < 0.5 * first and acc >= 0.95, (first, last, acc)


def state_cases(oracle):
    out = []
    n, ng = 40, 4
    g = oracle.grid(n, n, ng)
    spec = ic_spec("orszag_tang", 5 / 3, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0], ot_rho=25 / (36 * np.pi),
                   ot_p=5 / (12 * np.pi), noise_amplitude=1e-3, noise_seed=3)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    oracle.fill_ghosts(g, (0, 0, 0, 0), (0, 0, 0, 0), planes)
    out.append(("orszag_tang", g, 5 / 3, 7, (0, 0, 0, 0), planes))
    spec = ic_spec("rotor", 1.4, center=[0.5, 0.5], b_uniform=[5 / np.sqrt(4 * np.pi), 0, 0], p_ambient=1.0)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    oracle.fill_ghosts(g, (1, 1, 1, 1), (1, 1, 1, 1), planes)
    out.append(("rotor", g, 1.4, 7, (1, 1, 1, 1), planes))
    gb = oracle.grid(64, 8, 3, 1.0, 8 / 64)
    spec = ic_spec("brio_wu", 2.0, left=[1, 1, 0, 0, 0, 0.75, 1, 0], right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
    st, planes = oracle.seed(gb, spec)
    assert st == 0
    oracle.fill_ghosts(gb, (1, 1, 0, 0), (1, 1, 0, 0), planes)
    out.append(("brio_wu", gb, 2.0, 5, (1, 1, 0, 0), planes))
    return out


def test_features_match_oracle_bitwise(oracle):
    torch = torch_cuda()
    for name, g, gamma, _, _, planes in state_cases(oracle):
        qg = qgrid(g.nx, g.ny, g.ng, g.lx, g.ly)
        u = to_dev(planes)
        out = torch.zeros((g.nx * g.ny, 100), dtype=torch.float32, device="cuda")
        check(lib().qmhd_cell_features(C.byref(qg), gamma, ptrs(u), vp(out), C.c_void_p(scratch(qg).data_ptr()),
                                       None))
        check(lib().qmhd_stream_synchronize(None))
        got = out.cpu().numpy()
        exp = oracle.cell_features(g, gamma, planes)
        assert np.array_equal(got, exp), (name, np.abs(got - exp).max())


def test_labels_match_oracle(oracle):
    torch = torch_cuda()
    for name, g, gamma, order, sides, planes in state_cases(oracle):
        qg = qgrid(g.nx, g.ny, g.ng, g.lx, g.ly)
        u = to_dev(planes)
        out = torch.zeros(g.nx * g.ny, dtype=torch.uint8, device="cuda")
        fl = (C.c_int32 * 4)(*sides)
        check(lib().qmhd_cell_labels(C.byref(qg), order, fl, gamma, ptrs(u), 1e-6, vp(out),
                                     C.c_void_p(scratch(qg).data_ptr()), None))
        check(lib().qmhd_stream_synchronize(None))
        got = out.cpu().numpy()
        exp = oracle.cell_labels(g, order, sides, gamma, planes, 1e-6)
        agree = ((got != 0) == (exp != 0)).mean()
        assert agree >= 0.995, (name, agree)
        assert np.array_equal(got & 6, exp & 6), name  # fallbacks are not borderline here
        assert (exp != 0).any(), name


def test_indicator_kernel_matches_network_on_features(oracle):
    """The solver's fused indicator (features in LDS, MFMA network) flags
    exactly the interior cells whose logit (feature kernel + forward kernel)
    reaches the threshold; cells without a full stencil are flagged."""
    torch = torch_cuda()
    params = (rng(15).normal(0, 0.25, P)).astype(np.float32)
    for name, g, gamma, _, _, planes in state_cases(oracle)[:2]:
        qg = qgrid(g.nx, g.ny, g.ng, g.lx, g.ly)
        u = to_dev(planes)
        feats = torch.zeros((g.nx * g.ny, 100), dtype=torch.float32, device="cuda")
        sc = scratch(qg)
        check(lib().qmhd_cell_features(C.byref(qg), gamma, ptrs(u), vp(feats), C.c_void_p(sc.data_ptr()), None))
        check(lib().qmhd_stream_synchronize(None))
        z = gpu_logits(params, feats.cpu().numpy()).reshape(g.ny, g.nx)
        thr = float(np.median(z))
        p = torch.from_numpy(params).cuda()
        packed = packed_of(p)
        pitch, height = g.nx + 2 * g.ng, g.ny + 2 * g.ng
        flags = torch.full((pitch * height,), 7, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        check(lib().qmhd_cell_indicator(C.byref(qg), gamma, ptrs(u), vp(packed), vp(p), thr, vp(flags),
                                        C.c_void_p(sc.data_ptr()), None))
        check(lib().qmhd_stream_synchronize(None))
        f = flags.cpu().numpy().reshape(height, pitch)
        inner = f[g.ng:g.ng + g.ny, g.ng:g.ng + g.nx]
        assert np.array_equal(inner, (z >= thr).astype(np.uint8)), name
        assert np.all(f[:2, :] == 1) and np.all(f[-2:, :] == 1) and np.all(f[:, :2] == 1) and np.all(f[:, -2:] == 1)
        assert set(np.unique(f).tolist()) <= {0, 1}


def seeded_planes(oracle, n, kind):
    """gamma, boundary kind and the oracle-seeded state planes of a square
    n x n MP7 test problem."""
    g = oracle.grid(n, n, 4)
    if kind == "orszag_tang":
        gamma, sides = 5 / 3, "periodic"
        spec = ic_spec("orszag_tang", gamma, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0], ot_rho=25 / (36 * np.pi),
                       ot_p=5 / (12 * np.pi), noise_amplitude=1e-3, noise_seed=1)
    else:
        gamma, sides = 1.4, "outflow"
        spec = ic_spec("rotor", gamma, center=[0.5, 0.5], b_uniform=[5 / np.sqrt(4 * np.pi), 0, 0], p_ambient=1.0)
    st, planes = oracle.seed(g, spec)
    assert st == 0
    return gamma, sides, planes


def learned_pair(oracle, n, arithmetic, threshold, params=None, kind="orszag_tang", steps=3):
    from quasar_amd import mhd
    ng = 4
    gamma, sides, planes = seeded_planes(oracle, n, kind)
    sols = []
    for ind in ("none", "mlp"):
        cfg = mhd.MhdConfig()
        cfg.grid = mhd.Grid2D(n, n, 1.0, 1.0, 0.0, 0.0, ng)
        cfg.gamma = gamma
        cfg.reconstruction = "mp7"
        cfg.arithmetic = arithmetic
        cfg.indicator = ind
        cfg.boundary.set_fluid_all(sides)
        cfg.boundary.set_field_all(sides)
        s = mhd.MhdSolver2D(cfg)
        for k, nm in enumerate(mhd.STATE_COMPONENTS):
            s.seed_state(nm, planes[k])
        if ind == "mlp":
            s.set_indicator(params if params is not None else np.zeros(P, np.float32), threshold)
        sols.append(s)
    return sols


@pytest.mark.parametrize("arithmetic", ["fp64", "fp32"])
def test_learned_solver_with_every_cell_flagged_is_the_plain_method(oracle, arithmetic):
    plain, learned = learned_pair(oracle, 48, arithmetic, -np.inf)
    assert plain.run(4)[0] == 4 and learned.run(4)[0] == 4
    for nm in ("rho", "mx", "energy", "bx_face", "by_face"):
        np.testing.assert_array_equal(learned.state_component_to_host(nm), plain.state_component_to_host(nm))
    pf, total = learned.plain_faces()
    assert pf == total > 0


def test_learned_solver_with_no_cell_flagged_takes_the_cheap_path(oracle):
    """Smooth Orszag-Tang with every interior cell unflagged: linear
    fifth-order faces and HLL fluxes everywhere; the run stays close to the
    plain FP64 method (on this coarse 128^2 grid the two schemes differ by
    ~1e-4 in relative L1 after 10 steps) and div B stays at round-off."""
    from gpu_helpers import interior, rel_l1_state
    n = 128
    plain, learned = learned_pair(oracle, n, "fp64", np.inf)
    assert plain.run(10)[0] == 10 and learned.run(10)[0] == 10
    pf, total = learned.plain_faces()
    assert pf < 0.1 * total
    names = ("rho", "mx", "my", "mz", "energy", "bx_face", "by_face", "bz_cell")
    a = [interior(learned.state_component_to_host(nm), n, n, 4) for nm in names]
    b = [interior(plain.state_component_to_host(nm), n, n, 4) for nm in names]
    errs = [rel_l1_state(a, b, k) for k in range(8)]
    assert max(errs) <= 1e-3, errs
    assert errs[0] > 0
    assert learned.divergence_b_max() <= 1e-11


WEIGHTS = ROOT / "quasar_amd" / "data" / "indicator_mlp.npz"


@pytest.mark.skipif(not WEIGHTS.exists(), reason="trained indicator weights not present")
def test_trained_indicator_gains_over_both_extremes(oracle):
    """Rotor 192^2, 30 steps, FP32 + learned with the trained network, against
    the plain FP64 run: the learned run sends most faces to the cheap path and
    stays within the FP32+learned error bar, while sending every face to the
    cheap path (no indicator) is far less accurate or breaks down."""
    from gpu_helpers import interior, rel_l1_state
    from quasar_amd import learned as lrn
    params, thr, _ = lrn.load(WEIGHTS)
    n, steps = 192, 30
    names = ("rho", "mx", "my", "mz", "energy", "bx_face", "by_face", "bz_cell")

    def run(threshold):
        plain, lr = learned_pair(oracle, n, "fp32", threshold, params, kind="rotor")
        assert plain.run(steps)[0] == steps
        ok = True
        try:
            lr.run(steps)
        except (RuntimeError, ValueError):
            ok = False
        pf, total = lr.plain_faces()
        if not ok:
            return None, pf / max(total, 1)
        a = [interior(lr.state_component_to_host(nm), n, n, 4) for nm in names]
        b = [interior(plain.state_component_to_host(nm), n, n, 4) for nm in names]
        return rel_l1_state(a, b, 0), pf / max(total, 1)

    err_learned, frac_learned = run(thr)
    err_cheap, _ = run(np.inf)
    print("rotor: learned rho L1", err_learned, "plain-face share", frac_learned, "all-cheap rho L1", err_cheap)
    assert err_learned is not None and err_learned <= 1e-3
    assert frac_learned <= 0.6
    assert err_cheap is None or err_cheap >= 3 * err_learned


@pytest.mark.skipif(not WEIGHTS.exists(), reason="trained indicator weights not present")
def test_learned_step_reads_no_stale_interface_states(oracle):
    """A cheap face stores its interface states only where a flagged face's
    transverse stencil reads them; elsewhere the interface buffer keeps
    whatever it held.  Rotor 96^2 with the trained network (flagged and
    unflagged faces side by side): three learned steps give the same bits
    from a fresh solver and from one whose buffer holds the faces of six
    earlier steps."""
    from quasar_amd import learned as lrn
    from quasar_amd import mhd
    params, thr, _ = lrn.load(WEIGHTS)
    n = 96
    _, fresh = learned_pair(oracle, n, "fp32", thr, params, kind="rotor")
    _, used = learned_pair(oracle, n, "fp32", thr, params, kind="rotor")
    _, _, planes = seeded_planes(oracle, n, "rotor")
    assert used.run(6)[0] == 6
    for k, nm in enumerate(mhd.STATE_COMPONENTS):
        used.seed_state(nm, planes[k])
    assert fresh.run(3)[0] == 3 and used.run(3)[0] == 3
    pf, total = fresh.plain_faces()
    assert 0 < pf < total
    for nm in mhd.STATE_COMPONENTS + ("bx_face", "by_face"):
        np.testing.assert_array_equal(used.state_component_to_host(nm), fresh.state_component_to_host(nm))
