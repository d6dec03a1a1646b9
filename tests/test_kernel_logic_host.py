"""The kernels' own stencil code, built for the host, against the oracle.

tests/host_kernels compiles quasar_amd/csrc/qmhd_stencil.hpp and
qmhd_numerics.hpp with hipcc for the CPU, so these checks run the exact
functions the gfx950 kernels call (index maps, branches, fallbacks) without a
GPU.  The parts whose decisions are ill-conditioned are pinned bit for bit:
the face -> cell collocation (correctly rounded rational stencils and one-sided
closures, mhd_staggering.hpp:59-236) and the characteristic eigensystem with
its 1024-epsilon biorthogonality test (mhd_eigensystem.hpp:58-431).  The
reconstruction is compared per face at 1e-13 relative: the characteristic
projections use fused multiply-adds, so values differ by a few ulp, but every
branch (MP clip, MUSCL / piecewise-constant fallback, identity basis) must be
the oracle's.
"""
import numpy as np
import pytest

from host_kernels_ctypes import HostKernels
from oracle_ctypes import ic_spec

G53 = 5.0 / 3.0


@pytest.fixture(scope="module")
def hk():
    return HostKernels()


@pytest.mark.parametrize("ng,nx,ny,data", [(4, 13, 9, "normal"), (3, 11, 7, "normal"), (2, 9, 6, "normal"),
                                           (4, 3, 5, "normal"), (4, 11, 7, "integers"), (3, 11, 7, "dyadic"),
                                           (4, 11, 7, "thirds"), (2, 11, 7, "tiny")])
def test_collocation_bitwise(oracle, hk, ng, nx, ny, data):
    """Correctly rounded collocation, including the certified fast path of the
    exact sum: integer and dyadic samples make exact rounding ties, repeated
    thirds make near-ties, subnormal-range samples make the bound vanish."""
    rng = np.random.default_rng(ng * 100 + nx)
    g = oracle.grid(nx, ny, ng)
    npad = (nx + 2 * ng) * (ny + 2 * ng)
    if data == "integers":
        face = rng.integers(-1000, 1000, npad).astype(float)
    elif data == "dyadic":
        face = rng.integers(-8, 8, npad) * 2.0 ** rng.integers(-30, 30, npad)
    elif data == "thirds":
        face = np.full(npad, 1.0 / 3.0)
        face[rng.integers(0, npad, 7)] = 2.0 / 3.0
    elif data == "tiny":
        face = rng.normal(size=npad) * 1e-300
    else:
        face = rng.normal(size=npad) * 10.0 ** rng.integers(-3, 3, size=npad)
        face[rng.random(npad) < 0.1] = 0.0
        face[:5] = 1.0 / 3.0
    for axis in (0, 1):
        ghost = hk.collocate_ghost(g, axis, face)
        for co in (0, 1):
            got = hk.collocate(g, axis, co, face)
            if co == 0:
                np.testing.assert_array_equal(ghost, got)
            for j in range(-ng, ny + ng):
                for i in range(-ng, nx + ng):
                    k = (i + ng) + (nx + 2 * ng) * (j + ng)
                    ref = oracle.cell_bx(g, face, i, j, co) if axis == 0 else oracle.cell_by(g, face, i, j, co)
                    assert got[k] == ref, (axis, co, i, j, got[k], ref)


def _random_states(rng, n):
    out = []
    for k in range(n):
        rho = 10.0 ** rng.uniform(-1, 1)
        v = rng.uniform(-3, 3, 3)
        b = rng.uniform(-3, 3, 3)
        p = 10.0 ** rng.uniform(-2, 1)
        kind = k % 5
        if kind == 1:
            b[1] = b[2] = 0.0  # zero transverse field
        elif kind == 2:
            b[0] = 0.0  # zero normal field
        elif kind == 3:
            b[1] = b[2] = 0.0  # a ~ c_ax with zero transverse field: degenerate fans
            p = b[0] ** 2 / 1.4 * (1.0 + 10.0 ** rng.uniform(-14, -6))
        elif kind == 4:
            b[2] = 0.0
            b[1] = rng.uniform(-1e-7, 1e-7)
        e = p / 0.4 + 0.5 * rho * (v @ v) + 0.5 * (b @ b)
        out.append([rho, rho * v[0], rho * v[1], rho * v[2], e, b[0], b[1], b[2]])
    return out


def test_eigensystem_bitwise(oracle, hk):
    rng = np.random.default_rng(7)
    fallbacks = 0
    for s in _random_states(rng, 3000):
        Lr, Rr, _, fb = oracle.eigen_build(s, 1.4)
        L, R = hk.eigen(s, 1.4)
        np.testing.assert_array_equal(L, Lr)
        np.testing.assert_array_equal(R, Rr)
        fallbacks += int(fb)
    assert fallbacks > 0  # the degenerate family exercises the identity basis


def _max_face_rel(got, ref):
    rel = np.zeros_like(ref[0])
    for c in range(16):
        ok = np.isfinite(ref[c])
        assert np.array_equal(ok, np.isfinite(got[c]))
        d = np.abs(np.where(ok, got[c] - ref[c], 0.0)) / np.maximum(1.0, np.abs(np.where(ok, ref[c], 0.0)))
        rel = np.maximum(rel, d)
    return rel.max()


def _case(oracle, kind):
    if kind == "ot":
        n, ng, gam, F = 32, 4, G53, (0, 0, 0, 0)
        spec = ic_spec("orszag_tang", gam, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0], ot_rho=25 / (36 * np.pi),
                       ot_p=5 / (12 * np.pi), noise_amplitude=1e-3, noise_seed=3)
        g = oracle.grid(n, n, ng)
    elif kind == "rotor":
        n, ng, gam, F = 40, 4, 1.4, (1, 1, 1, 1)
        spec = ic_spec("rotor", gam, center=[0.5, 0.5], b_uniform=[5 / np.sqrt(4 * np.pi), 0, 0], p_ambient=1.0)
        g = oracle.grid(n, n, ng)
    else:
        nx, ny, ng, gam, F = 64, 4, 3, 2.0, (1, 1, 0, 0)
        spec = ic_spec("brio_wu", gam, left=[1, 1, 0, 0, 0, 0.75, 1, 0], right=[0.125, 0.1, 0, 0, 0, 0.75, -1, 0])
        g = oracle.grid(nx, ny, ng, 1.0, ny / nx)
    st, u = oracle.seed(g, spec)
    assert st == 0
    oracle.fill_ghosts(g, F, F, u)
    return g, gam, F, u


@pytest.mark.parametrize("kind,order", [("ot", 7), ("ot", 5), ("ot", 2), ("ot", 1), ("rotor", 7), ("rotor", 5),
                                        ("rotor", 2), ("brio", 5), ("brio", 7), ("brio", 2)])
def test_reconstruction_branches_match_oracle(oracle, hk, kind, order):
    g, gam, F, u = _case(oracle, kind)
    if order > 2 and g.ng < (order + 1) // 2:
        pytest.skip("halo too small for the order")
    for d in (0, 1):
        ref = [np.zeros_like(u[0]) for _ in range(16)]
        oracle.reconstruct(g, d, order, F, gam, u, ref)
        got = hk.reconstruct(g, d, order, F, gam, u)
        assert _max_face_rel(got, ref) <= 1e-13, (kind, order, d)
        if order <= 2:
            # MUSCL and piecewise-constant states round as the reference's (the
            # corner reads their exact L == R equalities).
            for c in range(16):
                np.testing.assert_array_equal(got[c], ref[c])


def _stage_states(oracle, g, order, flags, gamma, u0, steps):
    """Live and intermediate SSP-RK3 stage states of `steps` oracle steps
    composed from the launch-level operations (the solver's own sequence)."""
    from oracle_ctypes import OracleSolver
    solver = OracleSolver(oracle, g, gamma, 0.4, order, fluid=flags)
    solver.seed(u0)
    states = []
    for _ in range(steps):
        states.append(solver.state())
        solver.step_unchecked(solver.cfl_limit())
    return states


def _faces(oracle, g, order, flags, gamma, u):
    ifs, fls = [], []
    for d in (0, 1):
        r16 = [np.zeros_like(u[0]) for _ in range(16)]
        oracle.reconstruct(g, d, order, flags, gamma, u, r16)
        f8 = [np.zeros_like(u[0]) for _ in range(8)]
        oracle.face_flux(g, d, order, 1 if order <= 1 else 0, flags, gamma, r16, f8)
        ifs.append(r16)
        fls.append(f8)
    eq = []
    for d in (0, 1):
        e = np.ones(len(u[0]), dtype=np.int32)
        for c in range(8):
            e &= (ifs[d][c] == ifs[d][8 + c]).astype(np.int32)
        eq.append(e)
    return ifs, fls, eq


@pytest.mark.parametrize("flags", [(0, 0, 0, 0), (2, 2, 0, 0), (1, 1, 1, 1)])
def test_ct_corner_muscl_blast_matches_oracle(oracle, hk, flags):
    """Cell E_z bit for bit and the corner E_z (with its exact-equality 1-D
    shortcuts) on the states of a MUSCL blast, whose corners sit next to the
    shortcut decisions."""
    n, ng, gam = 32, 2, 5 / 3
    g = oracle.grid(n, n, ng)
    spec = ic_spec("blast", gam, rho_ambient=1.0, p_ambient=0.1, p_core=10.0, r_in=0.1, center=[0.5, 0.5],
                   b_uniform=[0.0, 0.5, 0.3])
    st, u0 = oracle.seed(g, spec)
    assert st == 0
    P = n + 2 * ng
    mask = np.zeros((P, P), bool)
    mask[ng:ng + n + 1, ng:ng + n + 1] = True
    mask = mask.ravel()
    for u in _stage_states(oracle, g, 2, flags, gam, u0, 4):
        oracle.fill_ghosts(g, flags, flags, u)
        ifs, fls, eq = _faces(oracle, g, 2, flags, gam, u)
        ref = oracle.ct_emf(g, 2, 0, flags, gam, u, ifs[0], ifs[1])
        bxc, byc = hk.collocate(g, 0, 0, u[5]), hk.collocate(g, 1, 0, u[6])
        cez = hk.cell_ez_lo(g, u, bxc, byc)
        phys = np.zeros((P, P), bool)
        phys[ng:ng + n + (1 if flags[3] not in (0, 4) else 0), ng:ng + n + (1 if flags[1] not in (0, 4) else 0)] = True
        phys = phys.ravel()
        np.testing.assert_array_equal(cez[phys], ref["cell_ez"][phys])
        ez = hk.corner(g, 2, flags, fls[0], fls[1], ref["cell_ez"], eq[0], eq[1])
        assert np.abs(ez - ref["ez"])[mask].max() <= 1e-15 * max(1.0, np.abs(ref["ez"]).max())


@pytest.mark.parametrize("order,ng", [(7, 4), (5, 3)])
def test_ct_corner_high_order_matches_oracle(oracle, hk, order, ng):
    """Corner E_z of the MP5/MP7 interpolations given the oracle's cell E_z:
    the stencils, index maps and shortcut decisions of the corner kernel."""
    n, gam = 24, G53
    g = oracle.grid(n, n, ng)
    spec = ic_spec("orszag_tang", gam, b_uniform=[1 / np.sqrt(4 * np.pi), 0, 0], ot_rho=25 / (36 * np.pi),
                   ot_p=5 / (12 * np.pi), noise_amplitude=1e-3, noise_seed=5)
    st, u = oracle.seed(g, spec)
    assert st == 0
    flags = (0, 0, 0, 0)
    oracle.fill_ghosts(g, flags, flags, u)
    ifs, fls, eq = _faces(oracle, g, order, flags, gam, u)
    ref = oracle.ct_emf(g, order, 0, flags, gam, u, ifs[0], ifs[1])
    ez = hk.corner(g, order, flags, fls[0], fls[1], ref["cell_ez"], eq[0], eq[1])
    P = n + 2 * ng
    mask = np.zeros((P, P), bool)
    mask[ng:ng + n + 1, ng:ng + n + 1] = True
    mask = mask.ravel()
    assert np.abs(ez - ref["ez"])[mask].max() <= 1e-14 * np.abs(ref["ez"]).max()


@pytest.mark.parametrize("kind,order", [("ot", 7), ("ot", 5), ("rotor", 7), ("rotor", 5), ("brio", 5)])
def test_cell_ez_tensor_recovery_matches_oracle(oracle, hk, kind, order):
    """The cell-EMF kernel's separable 4x4 (3x3) tensor recovery, EOS and Gauss
    combine against the oracle's cell-average E_z, per cell."""
    g, gam, F, u = _case(oracle, kind)
    ifs, _, _ = _faces(oracle, g, order, F, gam, u)
    ref = oracle.ct_emf(g, order, 0, F, gam, u, ifs[0], ifs[1])
    bxc, byc = hk.collocate(g, 0, 0, u[5]), hk.collocate(g, 1, 0, u[6])
    got = hk.cell_ez_tensor(g, order, gam, u, bxc, byc)
    nx, ny, ng = g.nx, g.ny, g.ng
    hx = 1 if F[1] not in (0, 4) else 0
    hy = 1 if F[3] not in (0, 4) else 0
    mask = np.zeros((ny + 2 * ng, nx + 2 * ng), bool)
    mask[ng:ng + ny + hy, ng:ng + nx + hx] = True
    mask = mask.ravel()
    scale = np.abs(ref["cell_ez"][mask]).max()
    assert np.abs(got - ref["cell_ez"])[mask].max() <= 1e-14 * scale
    # The corner's 1-D shortcuts read exact equalities of cell E_z along runs of
    # 2H+2 cells; with the kernel's own cell E_z they must decide as the
    # oracle's, so the corner E_z agrees to round-off.
    _, fls, eq = _faces(oracle, g, order, F, gam, u)
    ez = hk.corner(g, order, F, fls[0], fls[1], got, eq[0], eq[1])
    cmask = np.zeros((ny + 2 * ng, nx + 2 * ng), bool)
    cmask[ng:ng + ny + 1, ng:ng + nx + 1] = True
    cmask = cmask.ravel()
    assert np.abs(ez - ref["ez"])[cmask].max() <= 1e-14 * np.abs(ref["ez"][cmask]).max()


@pytest.mark.parametrize("d", [0, 1])
@pytest.mark.parametrize("density", [0.002, 0.02, 0.2])
def test_cheap_faces_store_every_state_a_flagged_face_reads(oracle, hk, d, density):
    """The learned form's cheap reconstruction stores a face's interface
    states only when face_states_read_back says so.  Brute force over the read
    pattern of the plain flux (flux_face: the 7 faces at offsets -3..3 along
    the face line, the same normal index): every face a flagged face in
    [0,nx] x [0,ny] reads must be marked, on random flag fields, edges and
    periodic high faces included.  Faces no flagged face reads are left
    unmarked (the predicate is the exact 7-face window, not a blanket)."""
    nx, ny, ng = 23, 19, 4
    g = oracle.grid(nx, ny, ng)
    rng = np.random.default_rng(int(density * 1000) + d)
    px, py = nx + 2 * ng, ny + 2 * ng
    flags = (rng.random(px * py) < density).astype(np.uint8)
    flagged, read_back = hk.face_read_back(g, d, flags)

    def k(i, j):
        return (i + ng) + px * (j + ng)

    if d == 0:
        faces = [(i, j) for j in range(-ng, ny + ng) for i in range(0, nx + 1)]
    else:
        faces = [(i, j) for j in range(0, ny + 1) for i in range(-ng, nx + ng)]
    # Flagged test restated: either adjacent cell flagged.
    for i, j in faces:
        a, b = (i - 1, j) if d == 0 else (i, j - 1)
        assert flagged[k(i, j)] == (flags[k(i, j)] | flags[k(a, b)] != 0)
    needed = set()
    for i, j in faces:
        if not (0 <= i <= nx and 0 <= j <= ny) or not flagged[k(i, j)]:
            continue
        for t in range(-3, 4):
            needed.add((i, j + t) if d == 0 else (i + t, j))
    face_set = set(faces)
    assert needed <= face_set
    for f in needed:
        assert read_back[k(*f)] == 1, f
    # Exactly the window: a face is marked iff a flagged face of the padded
    # range lies within 3 along its line.
    for i, j in faces:
        near = any(flagged[k(*f)] for f in ((i, j + t) if d == 0 else (i + t, j) for t in range(-3, 4))
                   if f in face_set)
        assert read_back[k(i, j)] == (1 if near else 0), (i, j)
    if density == 0.002:
        assert read_back.sum() < len(faces) // 2
