"""Pins for the CPU oracle (oracle/couplb_oracle.c) before it is trusted.

1. Golden fixtures: tests/golden/*.npz hold outputs of the reference's own
   numerics headers (oracle/_ref, see tests/golden/make_golden.py); the
   oracle must reproduce them bit for bit.
2. Mathematics the reference's algorithm must satisfy, chosen so a dropped
   term, a wrong sign/index or a transposed operand fails: lattice isotropy,
   delta-kernel point values and moments, BGK+Guo mass and momentum
   identities, streaming as an exact permutation, the closed-form Poiseuille
   and Couette profiles of validate.py, IBM exactness for linear fields,
   force conservation and interpolation/spreading adjointness.
3. The checkpoint header validation of couplb_io.h:721-811.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

import oracle
from tests.golden.make_golden import CASES, run_case
from tests.scenarios import run_oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# -------------------------------------------------------------------------
# 1. golden fixtures (reference headers) -- bit-exact
# -------------------------------------------------------------------------
@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_reproduces_reference_fixture_bit_exact(name):
    gold = np.load(os.path.join(GOLD, f"{name}.npz"))
    out = run_case(name, CASES[name], "oracle")
    for k in ("f", "rho", "ux", "uy", "uz", "atom_forces"):
        if k in gold:
            assert np.array_equal(out[k], gold[k]), k
    # diagnostics are OpenMP reductions: order-dependent, compare to 1e-13
    np.testing.assert_allclose(out["diag"], gold["diag"], rtol=1e-13, atol=1e-18)


def test_delta_kernels_match_reference_table():
    g = np.load(os.path.join(GOLD, "delta_kernels.npz"))
    roma = np.array([oracle.delta(oracle.ROMA, r) for r in g["r"]])
    pk = np.array([oracle.delta(oracle.PESKIN4, r) for r in g["r"]])
    assert np.array_equal(roma, g["roma"]) and np.array_equal(pk, g["peskin4"])


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("walls", [((0, 0), (1, 1), (0, 0)), ((4, 2), (3, 0), (1, 2)),
                                   ((2, 3), (4, 1), (3, 4))])
def test_oracle_matches_reference_headers_live(walls):
    rng = np.random.default_rng(sum(map(sum, walls)))
    n = 20
    x = rng.uniform(-0.5, 9.5, (n, 3))
    v = rng.normal(0, 2e-3, (n, 3))
    m = rng.uniform(0.2, 2.0, n)
    w = " ".join(f"wall_{ax} {walls[d][0]} {walls[d][1]}" for d, ax in enumerate("xyz"))
    cmd = (f"flow all couplb 9 8 10 0.07 1.0 {w} wall_vel 0.01 -0.004 0.002 "
           "gravity 3e-6 0 1e-6 kernel peskin4 md_per_lb 3 check_every 2")

    def atoms(step):
        return dict(x=x + 0.02 * step, v=v, f=np.zeros((n, 3)), mask=np.ones(n, np.int32), mass=m)

    a, fa = run_oracle(cmd, 12, dt=1.0 / 3.0, atoms_fn=atoms, numerics="oracle")
    b, fb = run_oracle(cmd, 12, dt=1.0 / 3.0, atoms_fn=atoms, numerics="ref")
    assert np.array_equal(a.f(), b.f())
    assert all(np.array_equal(p, q) for p, q in zip(fa, fb))


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
def test_density_clamp_and_non_finite_values_match_reference_headers():
    # couplb_lattice.h:178-186 clamps rho <= 1e-10 and non-finite rho; the
    # moment sums keep every product, so 0 * Inf = NaN spreads as it does
    # in the reference
    from tests.scenarios import bad_density_state
    n = (9, 7, 6)
    walls = ((0, 0), (1, 2), (0, 0))
    face = [[[0.0] * 3, [0.0] * 3] for _ in range(3)]
    face[1][1] = [0.02, 0.0, -0.01]
    f0 = bad_density_state(n)
    grids = [oracle.Grid(3, n, tau=0.7, walls=walls, face_vel=face, numerics=k) for k in ("oracle", "ref")]
    for g in grids:
        g.field(oracle.F_F)[:] = f0.reshape(-1)
    for step in range(2):
        clamps = []
        for g in grids:
            o = g.ops
            clamps.append(o.compute_macroscopic(g.g, 1))
            o.collide(g.g)
            o.exchange(g.g)
            o.stream(g.g)
        a, b = (g.field(oracle.F_F) for g in grids)
        assert np.array_equal(a, b, equal_nan=True)
        assert clamps[0] == clamps[1] >= 5
        assert np.isnan(a).any() and np.isfinite(a).any()


# -------------------------------------------------------------------------
# 2. mathematics
# -------------------------------------------------------------------------
@pytest.mark.parametrize("D", [2, 3])
def test_lattice_isotropy(D):
    e, w = oracle.lattice(D)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None] * e).sum(0), 0.0, atol=1e-16)
    M2 = np.einsum("q,qa,qb->ab", w, e, e)[:D, :D]
    assert np.allclose(M2, np.eye(D) / 3.0, atol=1e-15)
    M4 = np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e)[:D, :D, :D, :D]
    I = np.eye(D)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I)
           + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(M4, iso, atol=1e-15)


def test_delta_point_values_and_moments():
    d = oracle.delta
    assert d(oracle.ROMA, 0.0) == pytest.approx(2 / 3, abs=1e-15)
    assert d(oracle.ROMA, 1.0) == pytest.approx(1 / 6, abs=1e-15)
    assert d(oracle.ROMA, 1.5) == 0.0 and d(oracle.ROMA, 1.6) == 0.0
    assert d(oracle.PESKIN4, 0.0) == pytest.approx(0.5, abs=1e-15)
    assert d(oracle.PESKIN4, 1.0) == pytest.approx(0.25, abs=1e-15)
    assert d(oracle.PESKIN4, 2.0) == 0.0
    rng = np.random.default_rng(0)
    for r in rng.uniform(0, 1, 50):
        for k in (oracle.ROMA, oracle.PESKIN4):
            pts = r - np.arange(-2, 4)
            vals = np.array([d(k, p) for p in pts])
            assert abs(vals.sum() - 1.0) < 1e-14          # partition of unity
            assert abs((pts * vals).sum()) < 1e-14        # first moment
            assert np.all(vals >= 0)
        # Peskin4 even/odd split: sum over even i = sum over odd i = 1/2
        vals = np.array([d(oracle.PESKIN4, r - i) for i in range(-2, 4)])
        assert abs(vals[0::2].sum() - 0.5) < 1e-14


def _grid(D, n, tau=0.8, walls=((0, 0), (0, 0), (0, 0)), vel=None):
    ops = oracle.ops("oracle")
    p = oracle.GridParams()
    p.D, p.nx, p.ny, p.nz, p.tau, p.rho_lb = D, n[0], n[1], n[2], tau, 1.0
    for d in range(3):
        for s in range(2):
            p.wall[d][s] = walls[d][s]
            for c in range(3):
                p.face_vel[d][s][c] = vel[c] if vel is not None else 0.0
    create = C.CFUNCTYPE(C.c_void_p, C.POINTER(oracle.GridParams))(ops.contents.create)
    g = create(C.byref(p))
    gd = (C.c_int * 3)()
    nt = C.c_long()
    q = C.c_int()
    ops.contents.dims(g, gd, C.byref(nt), C.byref(q))
    return ops, g, (gd[0], gd[1], gd[2]), nt.value, q.value


def _arr(ops, g, which, n):
    return np.ctypeslib.as_array(ops.contents.dfield(g, which), shape=(n,))


@pytest.mark.parametrize("D", [2, 3])
def test_collision_conserves_mass_and_adds_guo_momentum(D):
    ops, g, gd, nt, Q = _grid(D, (5, 4, 3))
    rng = np.random.default_rng(D)
    f = _arr(ops, g, oracle.F_F, Q * nt)
    f[:] = rng.uniform(0.02, 0.12, Q * nt)
    F = [_arr(ops, g, k, nt) for k in (oracle.F_FX, oracle.F_FY, oracle.F_FZ)]
    for a in F:
        a[:] = rng.normal(0, 1e-3, nt)
    if D == 2:
        F[2][:] = 0.0
    f0 = f.copy().reshape(Q, nt)
    ops.contents.compute_macroscopic(g, 1)
    ops.contents.collide(g)
    df = f.reshape(Q, nt) - f0
    e, _ = oracle.lattice(D)
    type_ = np.ctypeslib.as_array(ops.contents.ifield(g, oracle.I_TYPE), shape=(nt,))
    gx, gy, gz = gd
    idx = np.arange(nt).reshape(gz, gy, gx)
    inner = (idx[1:-1, 1:-1, 1:-1] if D == 3 else idx[:, 1:-1, 1:-1]).ravel()
    inner = inner[type_[inner] == 0]
    assert np.max(np.abs(df[:, inner].sum(0))) < 1e-16 * 50
    for a in range(D):
        mom = (e[:, a][:, None] * df[:, inner]).sum(0)
        assert np.max(np.abs(mom - F[a][inner])) < 1e-16 * 50


def test_equilibrium_is_fixed_point_without_force():
    ops, g, gd, nt, Q = _grid(3, (4, 4, 4))
    rng = np.random.default_rng(1)
    rho = rng.uniform(0.9, 1.1, nt)
    u = rng.normal(0, 0.03, (3, nt))
    f = _arr(ops, g, oracle.F_F, Q * nt).reshape(Q, nt)
    f[:] = oracle.feq(3, rho, u)
    f0 = f.copy()
    ops.contents.compute_macroscopic(g, 1)
    ops.contents.collide(g)
    assert np.max(np.abs(f - f0)) < 1e-16 * 20


@pytest.mark.parametrize("D", [2, 3])
def test_periodic_stream_is_exact_permutation(D):
    n = (5, 4, 3)
    ops, g, gd, nt, Q = _grid(D, n)
    rng = np.random.default_rng(7)
    f = _arr(ops, g, oracle.F_F, Q * nt)
    f[:] = rng.uniform(0, 1, Q * nt)
    gx, gy, gz = gd
    before = f.reshape(Q, gz, gy, gx).copy()
    inner = before[:, 1:-1, 1:-1, 1:-1] if D == 3 else before[:, :, 1:-1, 1:-1]
    ops.contents.exchange(g)
    ops.contents.stream(g)
    after = _arr(ops, g, oracle.F_F, Q * nt).reshape(Q, gz, gy, gx)
    inner_after = after[:, 1:-1, 1:-1, 1:-1] if D == 3 else after[:, :, 1:-1, 1:-1]
    e, _ = oracle.lattice(D)
    for q in range(Q):
        shift = (e[q, 2], e[q, 1], e[q, 0]) if D == 3 else (0, e[q, 1], e[q, 0])
        expect = np.roll(inner[q], shift, axis=(0, 1, 2))
        assert np.array_equal(inner_after[q], expect)


def test_bounce_back_box_conserves_mass():
    ops, g, gd, nt, Q = _grid(3, (6, 5, 4), walls=((1, 1), (1, 1), (1, 1)))
    rng = np.random.default_rng(2)
    f = _arr(ops, g, oracle.F_F, Q * nt)
    f[:] = rng.uniform(0.01, 0.1, Q * nt)
    gx, gy, gz = gd

    def mass():
        ff = _arr(ops, g, oracle.F_F, Q * nt).reshape(Q, gz, gy, gx)
        return ff[:, 1:-1, 1:-1, 1:-1].sum()
    m0 = mass()
    for _ in range(20):
        ops.contents.compute_macroscopic(g, 1)
        ops.contents.collide(g)
        ops.contents.exchange(g)
        ops.contents.stream(g)
    assert abs(mass() - m0) < 1e-13 * m0


def test_poiseuille_fixture_passes_validate_py_rule():
    # validate.py:75-111 on the reference-generated C1 fixture
    gold = np.load(os.path.join(GOLD, "c1_poiseuille_2d.npz"))
    ux = gold["ux"].reshape(34, 10)[1:-1, 1:-1].mean(1)
    y = np.arange(32.0)
    exact = 5e-6 / (2 * 0.1) * (y + 0.5) * (32 - 0.5 - y)
    rel = np.max(np.abs(ux - exact)) / exact.max()
    assert rel < 0.01
    assert 6.3e-3 < exact.max() < 6.5e-3   # u_max = g H^2 / (8 nu) = 6.4e-3


def test_couette_matches_closed_form():
    # validate.py:132-166: u = U (y + dx/2) / H, moving wall on top
    gold = np.load(os.path.join(GOLD, "couette_2d.npz"))
    ux = gold["ux"].reshape(18, 8)[1:-1, 1:-1].mean(1)
    y = np.arange(16.0)
    exact = 0.02 * (y + 0.5) / 16
    assert np.max(np.abs(ux - exact)) / 0.02 < 0.01


def _interp(ops, g, x, lo, kern):
    out = (C.c_double * 3)()
    lo_a = (C.c_double * 3)(*lo)
    ops.contents.interpolate(g, x[0], x[1], x[2], lo_a, 1.0, kern, out)
    return np.array(out[:])


@pytest.mark.parametrize("kern", [oracle.ROMA, oracle.PESKIN4])
def test_interpolation_exact_for_linear_field(kern):
    n = (12, 12, 12)
    ops, g, gd, nt, Q = _grid(3, n)
    gx, gy, gz = gd
    k, j, i = np.meshgrid(np.arange(gz), np.arange(gy), np.arange(gx), indexing="ij")
    # node (i,j,k) ghost-inclusive sits at x = lo + (i - 1)
    X, Y, Z = (i - 1.0), (j - 1.0), (k - 1.0)
    for arr, val in ((oracle.F_UX, 0.01 + 1e-3 * X - 2e-3 * Y),
                     (oracle.F_UY, -0.02 + 5e-4 * Z),
                     (oracle.F_UZ, 3e-3 * X + 1e-3 * Y + 2e-3 * Z)):
        _arr(ops, g, arr, nt)[:] = val.ravel()
    rng = np.random.default_rng(kern)
    for p in rng.uniform(3.0, 8.0, (20, 3)):
        u = _interp(ops, g, p, (0.0, 0.0, 0.0), kern)
        exact = [0.01 + 1e-3 * p[0] - 2e-3 * p[1], -0.02 + 5e-4 * p[2],
                 3e-3 * p[0] + 1e-3 * p[1] + 2e-3 * p[2]]
        assert np.allclose(u, exact, rtol=0, atol=1e-15)


@pytest.mark.parametrize("kern", [oracle.ROMA, oracle.PESKIN4])
def test_spread_conserves_force_across_periodic_edges(kern):
    n = (8, 7, 6)
    ops, g, gd, nt, Q = _grid(3, n)
    rng = np.random.default_rng(11 + kern)
    lo = (C.c_double * 3)(0.0, 0.0, 0.0)
    total = np.zeros(3)
    for p in rng.uniform(-0.4, 7.4, (25, 3)):
        F = rng.normal(0, 1, 3)
        ops.contents.spread(g, p[0], p[1], p[2], F[0], F[1], F[2], 1.0, lo, 1.0, 1.0, kern)
        total += F
    ops.contents.exchange_forces(g)
    gx, gy, gz = gd
    got = np.array([_arr(ops, g, w, nt).reshape(gz, gy, gx)[1:-1, 1:-1, 1:-1].sum()
                    for w in (oracle.F_FX, oracle.F_FY, oracle.F_FZ)])
    assert np.allclose(got, total, rtol=0, atol=1e-13)


@pytest.mark.parametrize("kern", [oracle.ROMA, oracle.PESKIN4])
def test_spread_is_adjoint_of_interpolation_in_bulk(kern):
    n = (10, 10, 10)
    ops, g, gd, nt, Q = _grid(3, n)
    rng = np.random.default_rng(5)
    u = [rng.normal(0, 1, nt) for _ in range(3)]
    for w, a in zip((oracle.F_UX, oracle.F_UY, oracle.F_UZ), u):
        _arr(ops, g, w, nt)[:] = a
    p = np.array([4.3, 5.7, 4.9])
    F = np.array([0.3, -1.2, 0.7])
    ui = _interp(ops, g, p, (0.0, 0.0, 0.0), kern)
    lo = (C.c_double * 3)(0.0, 0.0, 0.0)
    ops.contents.spread(g, p[0], p[1], p[2], F[0], F[1], F[2], 1.0, lo, 1.0, 1.0, kern)
    Fn = [_arr(ops, g, w, nt) for w in (oracle.F_FX, oracle.F_FY, oracle.F_FZ)]
    lhs = sum((Fn[a] * u[a]).sum() for a in range(3))
    assert abs(lhs - F @ ui) < 1e-13


def test_fix_momentum_identity_per_lb_interval():
    """Periodic box, no gravity: fluid momentum change over one LB step equals
    minus the IBM impulse given to the atoms over the previous cycle
    (Guo momentum identity + exact spread normalisation)."""
    rng = np.random.default_rng(9)
    n = 15
    x = rng.uniform(0, 10, (n, 3))
    v = rng.normal(0, 1e-2, (n, 3))
    m = np.ones(n)
    N = 4
    fx = oracle.Fix(10, 10, 10, 0.1, 1.0, md_per_lb=N, dt=0.25, kernel=oracle.PESKIN4)

    def atoms():
        return dict(x=x, v=v, f=np.zeros((n, 3)), mask=np.ones(n, np.int32), mass=m)

    def momentum():
        f = fx.f()
        e, _ = oracle.lattice(3)
        inner = f[:, 1:-1, 1:-1, 1:-1].reshape(19, -1)
        # e_q in {0, +-1}: exact products, correctly rounded sums
        return np.array([math.fsum((e[:, a][:, None] * inner).ravel()) for a in range(3)])

    a = atoms()
    fx.setup(a)
    impulse = a["f"].sum(0) * 0.25   # setup sub-step, spread scaled by 1/N
    for cycle in range(3):
        p0 = momentum()
        for s in range(N):
            a = atoms()
            fx.post_force(a)
            if s == 0:
                dp = momentum() - p0         # LB step used the previous impulses
                # rounding of ~19k stored populations of size 0.05 bounds this
                assert np.max(np.abs(dp + impulse)) <= 2e-13, dp + impulse
                impulse = np.zeros(3)
            impulse += a["f"].sum(0) * 0.25
            fx.end_of_step(cycle * N + s + 1)


# -------------------------------------------------------------------------
# 3. checkpoint format (couplb_io.h:653-828)
# -------------------------------------------------------------------------
def test_checkpoint_roundtrip_and_header_validation(tmp_path):
    rng = np.random.default_rng(4)
    dims, Q = (4, 3, 5), 19
    nf = Q * 6 * 5 * 7
    f = rng.normal(size=nf)
    fn = str(tmp_path / "c.0.clbk")
    assert oracle.write_checkpoint(fn, 3, Q, dims, dims, 123, f)
    step, g = oracle.read_checkpoint(fn, 3, Q, dims, dims, nf)
    assert step == 123 and np.array_equal(f, g)
    assert os.path.getsize(fn) == 60 + 8 * nf   # 4+4+8+4+4+3*12 header bytes
    raw = bytearray(open(fn, "rb").read())
    for off, val, msg in ((0, 0x11, "bad magic"), (4, 2, "unsupported version"),
                          (16, 2, "D/Q mismatch"), (24, 9, "local grid mismatch"),
                          (36, 1, "offset mismatch"), (48, 7, "global grid mismatch")):
        bad = bytearray(raw)
        bad[off] = val
        open(fn, "wb").write(bytes(bad))
        with pytest.raises(IOError, match=msg):
            oracle.read_checkpoint(fn, 3, Q, dims, dims, nf)
    open(fn, "wb").write(bytes(raw[:-8]))
    with pytest.raises(IOError, match="short read"):
        oracle.read_checkpoint(fn, 3, Q, dims, dims, nf)


def test_opposite_moving_walls_couette_closed_form():
    # per-face wall velocities (wall_vel_face): u = U_lo + (U_hi - U_lo)(y + 1/2)/H,
    # the linear Couette profile that halfway bounce-back reproduces exactly
    from tests.scenarios import run_oracle
    cmd = ("flow all couplb 4 16 1 0.1 1.0 wall_y 2 2 "
           "wall_vel_face y lo -0.04 0 0 wall_vel_face y hi 0.04 0 0")
    fx, _ = run_oracle(cmd, 20000, dimension=2)
    ux = fx.field(oracle.F_UX).reshape(18, 6)[1:-1, 1:-1]
    y = np.arange(16.0)
    exact = -0.04 + 0.08 * (y + 0.5) / 16
    assert np.max(np.abs(ux - exact[:, None])) < 1e-10
    uy = fx.field(oracle.F_UY).reshape(18, 6)[1:-1, 1:-1]
    assert np.max(np.abs(uy)) < 1e-12


def _uniform_grid(D, n, u0):
    g = oracle.Grid(D, n)
    e, w = oracle.lattice(D)
    f = g.field(oracle.F_F).reshape(len(w), -1)
    eu = e @ u0
    f[:] = (w * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * u0 @ u0))[:, None]
    return g


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("kern", [oracle.ROMA, oracle.PESKIN4])
def test_mdf_single_marker_closed_form(D, kern):
    # one marker in uniform flow u0 (rho = 1): U = u0 + c F / 2 with
    # c = prod_d sum_i delta(r_d - i)^2, so the fixed point is F = 2 (V - u0) / c
    u0 = np.array([0.02, -0.01, 0.005 if D == 3 else 0.0])
    g = _uniform_grid(D, (12, 11, 10), u0)
    x = np.array([[5.3, 6.71, 4.45 if D == 3 else 0.0]])
    V = np.array([[0.004, 0.001, -0.002 if D == 3 else 0.0]])
    F, it, s0, s1 = oracle.mdf_solve(g, x, V, (0, 0, 0), 1.0, kern, 1e-13, relax=1.0, max_iter=2000)
    c = 1.0
    for d in range(D):
        l = x[0, d] + 1.0                  # phys_to_local, couplb_ibm.h:49-55
        i0 = np.floor(l)
        c *= sum(oracle.delta(kern, l - (i0 - 1 + t)) ** 2 for t in range(4))
    exact = 2 * (V[0] - u0) / c
    assert s1 < 1e-13 and s0 > 1e-3
    np.testing.assert_allclose(F[0, :D], exact[:D], rtol=1e-10, atol=1e-14)


def test_mdf_first_iteration_is_the_penalty_force():
    # from F = 0, one iteration gives F = 2 relax (V - I(u*)): the reference's
    # penalty force m xi/dt_LB (v - u_f) with m xi/dt_LB = 2 relax
    rng = np.random.default_rng(4)
    g = _uniform_grid(3, (10, 9, 8), np.array([0.01, 0.0, -0.01]))
    f = g.field(oracle.F_F)
    f *= 1 + 0.01 * rng.standard_normal(f.shape)
    x = rng.uniform(1.0, 7.0, (6, 3))
    V = rng.normal(0, 1e-2, (6, 3))
    F, it, _, _ = oracle.mdf_solve(g, x, V, (0, 0, 0), 1.0, oracle.PESKIN4, 1e-30, relax=0.7,
                                   max_iter=2)
    assert it == 2     # the slip at F = 0, one update, the slip after it
    g.ops.compute_macroscopic(g.g, 0)
    g.ops.exchange_velocity(g.g)
    out = (C.c_double * 3)()
    lo = (C.c_double * 3)(0, 0, 0)
    for k in range(6):
        g.ops.interpolate(g.g, *x[k], lo, 1.0, oracle.PESKIN4, out)
        assert np.array_equal(F[k], 2 * 0.7 * (V[k] - np.array(out[:])))


def test_mdf_cg_and_richardson_reach_the_same_forces():
    # markers on a sphere in a sheared, walled box: both solvers converge to
    # the one fixed point; conjugate gradients in far fewer passes
    rng = np.random.default_rng(8)
    face = [[[0, 0, 0], [0, 0, 0]], [[-0.02, 0, 0], [0.02, 0, 0]], [[0, 0, 0], [0, 0, 0]]]
    res = {}
    for method in ("richardson", "cg"):
        g = oracle.Grid(3, (14, 12, 12), walls=((0, 0), (2, 2), (0, 0)), face_vel=face)
        e, w = oracle.lattice(3)
        f = g.field(oracle.F_F).reshape(19, -1)
        f[:] = w[:, None] * (1 + 0.002 * np.random.default_rng(3).standard_normal(f.shape))
        k = np.arange(40) + 0.5
        phi, th = np.arccos(1 - 2 * k / 40), np.pi * (1 + 5 ** 0.5) * k
        x = np.stack([6.6 + 2.2 * np.cos(th) * np.sin(phi), 5.9 + 2.2 * np.sin(th) * np.sin(phi),
                      6.1 + 2.2 * np.cos(phi)], 1)
        V = np.tile([0.003, 0.0, -0.001], (40, 1))
        res[method] = oracle.mdf_solve(g, x, V, (0, 0, 0), 1.0, oracle.PESKIN4, 1e-8,
                                       max_iter=8000, method=method, relax=1.5)
    (Fr, ir, _, sr), (Fc, ic, _, sc) = res["richardson"], res["cg"]
    assert sr < 1e-8 and sc < 1e-8
    assert ic < ir / 20
    # a slip tolerance bounds the force error only by tol / lambda_min(A); with
    # overlapping Peskin4 stencils A is ill-conditioned, so compare at 1e-3
    assert np.max(np.abs(Fr - Fc)) < 1e-3 * np.max(np.abs(Fc))
