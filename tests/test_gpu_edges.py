"""Edge cases of the device IBM entry points: empty marker sets, markers
that cannot be placed (non-finite or far outside the box), rows longer than
1024 nodes, node runs longer than the pair maps' run lengths hold, misuse of
the multi-direct-forcing call, and density clamps."""
import numpy as np
import pytest

from tests.scenarios import DeviceArray, bad_density_state

pytestmark = pytest.mark.gpu

KW = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0, inv_N=1.0,
          force_scale=1.0)


def test_empty_marker_set_is_a_no_op():
    import couplb_amd
    d = couplb_amd.Device(3, 8, 8, 8, tau=0.8)
    d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    f0 = d.get_state()
    empty = np.zeros((0, 3))
    d.ibm(empty, empty, np.zeros(0), np.zeros((0, 3)), **KW)
    st = d.mdf_device(0, None, None, None, None, (0, 0, 0), 1.0, 1.0, 1.0)
    assert st["iterations"] == 0
    d.step((0.0, 0.0, 0.0), 0)
    d.sync()
    # uniform equilibrium without force is a fixed point of the step
    assert np.allclose(d.get_state()[:, 1:-1, 1:-1, 1:-1], f0[:, 1:-1, 1:-1, 1:-1], rtol=0, atol=1e-15)


def test_unplaceable_markers_get_the_pure_damping_force():
    # no stencil weight: the interpolated velocity is 0 (couplb_ibm.h:88-89), so
    # F = m xi (0 - v), and nothing is spread
    import couplb_amd
    d = couplb_amd.Device(3, 8, 8, 8, tau=0.8, kernel=couplb_amd.PESKIN4)
    d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    x = np.array([[np.nan, 1.0, 1.0], [1e12, 2.0, 2.0], [3.0, np.inf, 3.0]])
    v = np.array([[1e-3, 2e-3, 3e-3], [0.0, -1e-3, 0.0], [5e-4, 0.0, 0.0]])
    m = np.array([1.0, 2.0, 3.0])
    f = np.zeros((3, 3))
    d.ibm(x, v, m, f, **KW)
    assert np.array_equal(f, -m[:, None] * v)
    f_before = d.get_state()
    d.step((0.0, 0.0, 0.0), 0)
    d.sync()
    g = couplb_amd.Device(3, 8, 8, 8, tau=0.8)
    g.set_state(f_before)
    g.step((0.0, 0.0, 0.0), 0)
    assert np.array_equal(d.get_state()[:, 1:-1, 1:-1, 1:-1], g.get_state()[:, 1:-1, 1:-1, 1:-1])


def test_rows_longer_than_1024_nodes_match_the_oracle():
    # the reference's IBM has no row-length limit (couplb_ibm.h:57-133);
    # markers at both x ends (periodic fold) and in the middle of 1100-node rows
    from tests.scenarios import atoms_from, interior, rel_l2, run_device, run_oracle
    rng = np.random.default_rng(8)
    xs = np.concatenate([rng.uniform(0.0, 3.0, 6), rng.uniform(540.0, 560.0, 6), rng.uniform(1097.0, 1100.0, 6)])
    x = np.stack([xs, rng.uniform(0.5, 5.5, 18), rng.uniform(0.5, 4.5, 18)], 1)
    base = atoms_from(x, rng.normal(0, 1e-3, (18, 3)), np.ones(18))

    def atoms(step):
        a = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}
        a["x"] = a["x"] + a["v"] * 0.5 * step
        return a

    cmd = "long all couplb 1100 6 5 0.1 1.0 md_per_lb 2 kernel peskin4 gravity 1e-5 0.0 0.0"
    print("case IBM on 1100-node rows", flush=True)
    dev, fd = run_device(cmd, 6, dimension=3, dt=0.5, atoms_fn=atoms)
    ref, fo = run_oracle(cmd, 6, dimension=3, dt=0.5, atoms_fn=atoms)
    a = interior(dev.device_state().reshape(19, 7, 8, 1102), 3)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    ferr = max(rel_l2(p, q) for p, q in zip(fd, fo))
    print(f"  rel-L2(f) {err:.3e} rel-L2(F_atom) {ferr:.3e}")
    assert err <= 1e-12 and ferr <= 1e-11


def test_node_runs_longer_than_a_run_length_holds_match_the_oracle():
    # 70,000 markers in one cell: each of its 64 stencil nodes gets 70,000
    # pairs, more than the pair maps' 16-bit run lengths hold; the fixed
    # markers' step then spreads through the node rows instead
    from tests.scenarios import atoms_from, interior, rel_l2, run_device, run_oracle
    rng = np.random.default_rng(12)
    n = 70000
    x = 5.5 + rng.uniform(-0.4, 0.4, (n, 3))
    base = atoms_from(x, rng.normal(0, 1e-3, (n, 3)), np.full(n, 1e-6))

    def atoms(step):
        return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in base.items()}

    cmd = "dense all couplb 10 10 10 0.1 1.0 kernel peskin4 gravity 1e-5 0.0 0.0"
    print("case 70,000 markers in one cell", flush=True)
    dev, fd = run_device(cmd, 4, dimension=3, dt=1.0, atoms_fn=atoms)
    ref, fo = run_oracle(cmd, 4, dimension=3, dt=1.0, atoms_fn=atoms)
    a = interior(dev.device_state().reshape(19, 12, 12, 12), 3)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    ferr = max(rel_l2(p, q) for p, q in zip(fd, fo))
    print(f"  rel-L2(f) {err:.3e} rel-L2(F_atom) {ferr:.3e}")
    assert err <= 1e-12 and ferr <= 1e-10


def test_mdf_must_be_the_only_ibm_call_of_an_interval():
    import couplb_amd
    d = couplb_amd.Device(3, 12, 12, 12, tau=0.8, kernel=couplb_amd.PESKIN4)
    d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    x = DeviceArray(np.array([[6.2, 6.1, 5.9], [5.1, 6.6, 6.3]]))
    V = DeviceArray(np.zeros((2, 3)))
    fb = DeviceArray(np.zeros((2, 3)))
    d.mdf_device(2, x.addr, V.addr, None, fb.addr, (0, 0, 0), 1.0, 1.0, 1.0, tol=1e-8)
    with pytest.raises(couplb_amd.CoupLBError, match="only IBM call"):
        d.mdf_device(2, x.addr, V.addr, None, fb.addr, (0, 0, 0), 1.0, 1.0, 1.0, tol=1e-8)
    with pytest.raises(couplb_amd.CoupLBError):
        d.mdf_device(2, x.addr, V.addr, None, fb.addr, (0, 0, 0), 1.0, 1.0, 1.0, tol=-1.0)
    d.step((0.0, 0.0, 0.0), 0)
    st = d.mdf_device(2, x.addr, V.addr, None, fb.addr, (0, 0, 0), 1.0, 1.0, 1.0, tol=1e-8,
                      positions_unchanged=True)
    assert st["converged"]


@pytest.mark.parametrize("walls", [((0, 0), (1, 2), (0, 0)), ((0, 0), (0, 0), (0, 0))])
def test_density_clamp_matches_the_oracle(walls):
    # couplb_lattice.h:178-186: rho <= 1e-10 and non-finite rho (NaN, +Inf)
    # are clamped to 1e-10 before the collision; the count comes from the
    # next compute_macroscopic(false) (fix_couplb.cpp:838)
    import couplb_amd
    import oracle
    from tests.scenarios import interior
    n = (9, 7, 6)
    face = [[[0.0] * 3, [0.0] * 3] for _ in range(3)]
    face[1][1] = [0.02, 0.0, -0.01]
    f0 = bad_density_state(n)
    print("case density clamp", walls, flush=True)
    g = oracle.Grid(3, n, tau=0.7, walls=walls, face_vel=face)
    g.field(oracle.F_F)[:] = f0.reshape(-1)
    d = couplb_amd.Device(3, *n, tau=0.7, walls=walls, face_vel=face)
    d.set_state(f0)
    o = g.ops
    for step in range(2):
        o.compute_macroscopic(g.g, 1)
        o.collide(g.g)
        o.exchange(g.g)
        o.stream(g.g)
        d.step((0.0, 0.0, 0.0), 0)
        fo = interior(g.field(oracle.F_F).reshape(19, n[2] + 2, n[1] + 2, n[0] + 2), 3)
        fd = interior(d.get_state(), 3)
        assert np.isnan(fo).any() and np.isfinite(fo).any()
        # FP64 steps without markers equal the reference bit for bit where
        # the values are finite, and are non-finite where the reference's
        # are (NaN or an infinity there: the collision skips the products
        # with zero lattice components, which turn an infinite velocity
        # into NaN in the reference)
        fin = np.isfinite(fo)
        assert np.array_equal(np.isfinite(fd), fin), (step, np.argwhere(np.isfinite(fd) != fin)[:5])
        assert np.array_equal(fd[fin], fo[fin]), step
        clamps_ref = o.compute_macroscopic(g.g, 0)
        assert clamps_ref >= 5
        assert d.diagnostics()["clamps"] == clamps_ref
