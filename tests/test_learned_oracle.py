"""CPU checks of the learned form's oracle: the network's loss gradient against
finite differences of the exact network, its forward pass against an
independent numpy evaluation (with and without BF16 operand rounding), the
cell features against closed forms, the HLL flux against the HLL of the
reference numerics with Davis speeds, and the labels on problems whose
troubled cells are known."""
import numpy as np
import pytest

from oracle_ctypes import ic_spec

P = 4321
W1, B1, W2, B2, W3, B3 = 0, 3200, 3232, 4256, 4288, 4320


def rng(seed):
    return np.random.Generator(np.random.Philox(key=seed))


def random_net(seed, scale=0.3):
    r = rng(seed)
    p = r.normal(0.0, scale, P).astype(np.float32)
    return p


def bf16(a):
    """Round-to-nearest-even to bfloat16, computed on the float32 bit pattern."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def numpy_forward(p, x, rounded):
    q = (lambda v: bf16(v).astype(np.float64)) if rounded else (lambda v: np.asarray(v, np.float64))
    w1 = p[W1:B1].reshape(100, 32)
    w2 = p[W2:B2].reshape(32, 32)
    xs = q(x)
    z1 = xs @ q(w1) + p[B1:W2]
    a1 = q(np.maximum(z1, 0).astype(np.float32)) if rounded else np.maximum(z1, 0)
    z2 = a1 @ q(w2) + p[B2:W3]
    return np.maximum(z2, 0) @ p[W3:B3].astype(np.float64) + float(p[B3])


def test_bf16_rounding_reference_points():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    assert bf16(one + np.float32(2.0 ** -8)) == one           # tie -> even (down)
    assert bf16(one + ulp + np.float32(2.0 ** -8)) == one + 2 * ulp  # tie -> even (up)
    assert bf16(np.float32(3.0)) == np.float32(3.0)
    assert bf16(np.float32(-1.00390625)) == np.float32(-1.0)


@pytest.mark.parametrize("rounded", [False, True])
def test_forward_matches_numpy(oracle, rounded):
    p = random_net(1)
    x = rng(2).uniform(-1, 1, (64, 100)).astype(np.float32)
    got = oracle.mlp_forward(p, x, bf16=rounded)
    exp = numpy_forward(p, x, rounded)
    # Rounding of a hidden activation that sits on a BF16 tie can differ
    # between float64 (numpy) and float32-then-round (oracle) evaluation; the
    # bulk must agree to round-off.
    err = np.abs(got - exp)
    assert np.median(err) <= 1e-12 * (1 + np.abs(exp).max())
    assert (err > 1e-3).sum() <= 1


def test_gradient_matches_finite_differences(oracle):
    """Exact network (no operand rounding): central differences in float64."""
    p = random_net(3)
    p64 = p.astype(np.float64)
    r = rng(4)
    x = r.uniform(-1, 1, (40, 100)).astype(np.float32)
    y = (r.uniform(0, 1, 40) < 0.4).astype(np.float32)
    w = r.uniform(0.5, 1.5, 40).astype(np.float32)
    loss, g = oracle.mlp_loss_grad(p, x, y, w, bf16=False, params64=p64)
    picks = np.concatenate([r.choice(np.arange(W1, B1), 30, replace=False), np.arange(B1, B1 + 8),
                            r.choice(np.arange(W2, B2), 20, replace=False), np.arange(B2, B2 + 8),
                            np.arange(W3, W3 + 8), [B3]])
    h = 1e-6
    for k in picks:
        pp = p64.copy()
        pp[k] += h
        lp, _ = oracle.mlp_loss_grad(p, x, y, w, bf16=False, params64=pp)
        pp[k] -= 2 * h
        lm, _ = oracle.mlp_loss_grad(p, x, y, w, bf16=False, params64=pp)
        num = (lp - lm) / (2 * h)
        assert abs(num - g[k]) <= 1e-6 * (1 + abs(g[k])), (k, num, g[k])
    assert loss > 0


def test_bf16_gradient_tracks_exact_gradient(oracle):
    p = random_net(5)
    r = rng(6)
    x = r.uniform(-1, 1, (128, 100)).astype(np.float32)
    y = (r.uniform(0, 1, 128) < 0.5).astype(np.float32)
    w = np.ones(128, np.float32)
    _, ge = oracle.mlp_loss_grad(p, x, y, w, bf16=False)
    _, gb = oracle.mlp_loss_grad(p, x, y, w, bf16=True)
    assert np.linalg.norm(gb - ge) <= 0.05 * np.linalg.norm(ge)


def padded_planes(n, ng, fn):
    c = (np.arange(-ng, n + ng) + 0.5) / n
    x, y = np.meshgrid(c, c)
    return [np.ascontiguousarray(a.ravel()) for a in fn(x, y)]


def test_features_of_uniform_state_vanish(oracle):
    n, ng = 8, 4
    g = oracle.grid(n, n, ng)
    one = lambda x, y: [np.full_like(x, v) for v in (1.3, 0.2, -0.1, 0.05, 3.0, 0.7, -0.4, 0.1)]
    f = oracle.cell_features(g, 5 / 3, padded_planes(n, ng, one))
    # Only the collocated field can carry round-off (its high-order
    # face->cell rule is exact for constants up to the last bit).
    assert np.all(f[:, :50] == 0.0)
    assert np.max(np.abs(f)) <= 1e-14


def test_density_ramp_features_closed_form(oracle):
    """rho = a + b x (v = 0, uniform p and B): the density features of a cell
    are b h dx / (rho_c + 2 b h); every other quantity is uniform except |v|
    (zero), so its features vanish too."""
    n, ng, a, b = 16, 4, 2.0, 0.5
    g = oracle.grid(n, n, ng)
    gamma, p0 = 1.4, 0.6

    def planes(x, y):
        rho = a + b * x
        z = np.zeros_like(x)
        return [rho, z, z, z, np.full_like(x, p0 / (gamma - 1)), z, z, z]

    f = oracle.cell_features(g, gamma, padded_planes(n, ng, planes)).reshape(n, n, 100)
    h = 1.0 / n
    for (i, j) in [(0, 0), (5, 7), (15, 15)]:
        rc = a + b * (i + 0.5) * h
        for dy in range(-2, 3):
            for dx in range(-2, 3):
                exp = (b * dx * h) / (rc + 2 * b * h)
                assert abs(f[j, i, 5 * (dy + 2) + (dx + 2)] - exp) <= 2e-7 * (1 + abs(exp))
        assert np.all(f[j, i, 25:] == 0.0)


def test_features_are_scale_free(oracle):
    """Scaling rho and p by s and B by sqrt(s) (velocities unchanged) leaves
    every feature unchanged up to rounding."""
    n, ng, gamma = 12, 4, 5 / 3

    def make(s):
        def planes(x, y):
            rho = s * (1 + 0.3 * np.sin(2 * np.pi * x) * np.cos(2 * np.pi * y))
            vx, vy = 0.2 * np.sin(2 * np.pi * y), 0.1 * np.cos(2 * np.pi * x)
            p = s * (0.5 + 0.1 * np.cos(2 * np.pi * (x + y)))
            bx = np.sqrt(s) * 0.3 * np.ones_like(x)
            by = np.sqrt(s) * 0.2 * np.sin(2 * np.pi * x)
            e = p / (gamma - 1) + 0.5 * rho * (vx ** 2 + vy ** 2) + 0.5 * (bx ** 2 + by ** 2)
            z = np.zeros_like(x)
            return [rho, rho * vx, rho * vy, z, e, bx, by, z]
        return planes

    g = oracle.grid(n, n, ng)
    f1 = oracle.cell_features(g, gamma, padded_planes(n, ng, make(1.0)))
    f2 = oracle.cell_features(g, gamma, padded_planes(n, ng, make(37.0)))
    assert np.max(np.abs(f1 - f2)) <= 1e-5


def test_hll_davis_is_hll_with_davis_speeds(oracle):
    r = rng(7)
    gamma = 5 / 3
    for _ in range(50):
        l = np.array([r.uniform(0.5, 2), r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1), 0,
                      0.4, r.uniform(-1, 1), r.uniform(-1, 1)])
        rr = np.array([r.uniform(0.5, 2), r.uniform(-1, 1), r.uniform(-1, 1), r.uniform(-1, 1), 0,
                       0.4, r.uniform(-1, 1), r.uniform(-1, 1)])
        for s in (l, rr):
            s[4] = r.uniform(1, 3) + 0.5 * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2) / s[0] + 0.5 * (s[5:] ** 2).sum()
        vl, vr = l[1] / l[0], rr[1] / rr[0]
        cf = max(oracle.fast_speed(l, 0, gamma), oracle.fast_speed(rr, 0, gamma))
        sl, sr = min(vl, vr) - cf, max(vl, vr) + cf
        exp = oracle.hll_flux_x(l, rr, sl, sr, gamma)
        got = oracle.hll_davis_flux_x(l, rr, gamma)
        assert np.allclose(got, exp, rtol=0, atol=1e-14 * (1 + np.abs(exp).max()))


def test_labels_mark_the_brio_wu_jump_only(oracle):
    nx, ny, ng = 64, 8, 3
    g = oracle.grid(nx, ny, ng, 1.0, ny / nx)
    spec = ic_spec("brio_wu", 2.0, left=[1, 1, 0, 0, 0, 0.75, 1, 0], right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
    st, planes = oracle.seed(g, spec)
    assert st == 0
    oracle.fill_ghosts(g, (1, 1, 0, 0), (1, 1, 0, 0), planes)
    bits = oracle.cell_labels(g, 5, (1, 1, 0, 0), 2.0, planes, 1e-6).reshape(ny, nx)
    cols = np.flatnonzero(bits.any(axis=0))
    # The jump sits between cells 31 and 32; MP5 stencils reach 3 cells.
    assert {31, 32} <= set(cols.tolist())
    assert cols.min() >= 31 - 3 and cols.max() <= 32 + 3
    assert np.all(bits[:, :25] == 0) and np.all(bits[:, 40:] == 0)
    # Every row is the same (the problem is one-dimensional).
    assert np.all(bits == bits[0])


def test_labels_of_smooth_orszag_tang_are_rare(oracle):
    n, ng = 64, 4
    g = oracle.grid(n, n, ng)
    spec = ic_spec("orszag_tang", 5 / 3, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0],
                   ot_rho=25 / (36 * np.pi), ot_p=5 / (12 * np.pi))
    st, planes = oracle.seed(g, spec)
    assert st == 0
    oracle.fill_ghosts(g, (0, 0, 0, 0), (0, 0, 0, 0), planes)
    bits = oracle.cell_labels(g, 7, (0, 0, 0, 0), 5 / 3, planes, 1e-6)
    assert (bits & 2).sum() == 0 and (bits & 4).sum() == 0  # no fallbacks in smooth data
    assert (bits != 0).mean() < 0.25
