"""MI355X path vs the CPU oracle on the same inputs (FP64: <= 1e-12 relative).

Each test runs one `fix couplb` deck through the HIP library (couplb_amd.Fix
-> include/couplb_fix.h -> include/couplb_hip.h) and through the oracle's
restatement of the reference, then compares the post-stream populations in
the reference's layout, the per-atom IBM forces and the diagnostics.
"""
import os

import numpy as np
import pytest

import oracle
from tests.golden.make_golden import CASES, ibm_atoms, moving
from tests.scenarios import (atoms_from, copy_atoms, interior, rel_l2, run_device,
                             run_oracle)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def test_c1_poiseuille_2d_matches_oracle_and_parabola(tmp_path):
    # D2Q9 8x32, nu=0.1, no-slip y walls, g=5e-6, rest start, 50000 steps
    out = tmp_path / "c1_profile.dat"
    cmd = ("flow all couplb 8 32 1 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 "
           f"check_every 2000 output 10000 {out}")
    print("case C1 D2Q9 8x32 50000 steps", flush=True)
    dev, _ = run_device(cmd, 50000, dimension=2)
    ref, _ = run_oracle(cmd, 50000, dimension=2)
    fd = interior(dev.device_state().reshape(9, 1, 34, 10), 2)
    fo = interior(ref.f(), 2)
    err = rel_l2(fd, fo)
    print("C1 rel-L2(f) =", err, "bit-exact", np.array_equal(fd, fo))
    assert np.array_equal(fd, fo)    # FP64 path without markers: identical bits
    dd, od = dev.last_diag(), ref.last_diag()
    assert abs(dd["mass"] - od["mass"]) <= 1e-12 * od["mass"]
    assert abs(dd["Ma"] - od["Ma"]) <= 1e-10 * od["Ma"]
    # validate.py:75-111: u(y) = g/(2 nu) (y + dx/2)(H - dx/2 - y), PASS < 1 %
    r, ux, uy = dev.last_profile()
    y = np.arange(32, dtype=float)
    exact = 5e-6 / (2 * 0.1) * (y + 0.5) * (32 - 0.5 - y)
    rel = np.max(np.abs(ux - exact)) / exact.max()
    print("C1 parabola max rel err", rel)
    assert rel < 0.01
    ro, uo, vo = ref.last_profile()
    assert np.max(np.abs(ux - uo)) <= 1e-12 * np.max(np.abs(uo))


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", sorted(CASES))
def test_device_matches_reference_fixture(name):
    """Replay each fixture written by the reference's own headers. Without
    markers the FP64 path is bit-identical; with markers only the spreading
    summation order differs (<= 1e-12 relative)."""
    c = CASES[name]
    gold = np.load(os.path.join(GOLD, f"{name}.npz"))
    atoms_fn = None
    if c["atoms"] is not None:
        n, L, seed = c["atoms"]
        base = ibm_atoms(n, L, seed)
        if c["dim"] == 2:
            base["x"][:, 2] = 0.0
            base["v"][:, 2] = 0.0
        atoms_fn = moving(base)
    print("case", name, flush=True)
    dev, forces = run_device(c["cmd"], c["steps"], dimension=c["dim"], dt=c["dt"], atoms_fn=atoms_fn)
    fd = interior(dev.device_state().reshape(gold["f"].shape), c["dim"])
    fo = interior(gold["f"], c["dim"])
    err = rel_l2(fd, fo)
    print(f"{name}: rel-L2(f) = {err:.3e} bit-exact={np.array_equal(fd, fo)}")
    if c["atoms"] is None:
        assert np.array_equal(fd, fo)
    else:
        assert err <= TOL
        ferr = max(rel_l2(a, b) for a, b in zip(forces, gold["atom_forces"]))
        assert ferr <= 1e-11


WALLS = [
    ((0, 0), (1, 1), (0, 0)),
    ((0, 0), (1, 2), (0, 0)),
    ((3, 1), (2, 4), (0, 0)),
    ((1, 3), (0, 0), (2, 3)),
    ((0, 0), (0, 0), (0, 0)),
    ((2, 2), (3, 3), (1, 4)),
]


def _ibm_atoms(n, L, seed):
    rng = np.random.default_rng(seed)
    x = rng.uniform(0.0, L, (n, 3))
    v = rng.normal(0.0, 1e-3, (n, 3))
    m = rng.uniform(0.5, 1.5, n)
    return atoms_from(x, v, m)


@pytest.mark.parametrize("walls", WALLS)
@pytest.mark.parametrize("kernel", ["roma", "peskin4"])
def test_3d_walls_ibm_matches_oracle(walls, kernel):
    w = " ".join(f"wall_{ax} {walls[d][0]} {walls[d][1]}" for d, ax in enumerate("xyz"))
    cmd = (f"flow all couplb 12 10 14 0.1 1.0 {w} wall_vel 0.01 0.0 0.005 "
           f"gravity 1e-5 2e-6 -1e-6 kernel {kernel} md_per_lb 2 check_every 3")
    print("case", cmd, flush=True)
    base = _ibm_atoms(40, 10.0, 7)

    def atoms_fn(step):
        a = copy_atoms(base)
        a["x"] = a["x"] + 0.05 * np.sin(0.3 * step + np.arange(len(a["mass"])))[:, None]
        return a

    dev, fdv = run_device(cmd, 24, dimension=3, dt=0.5, atoms_fn=atoms_fn)
    ref, fro = run_oracle(cmd, 24, dimension=3, dt=0.5, atoms_fn=atoms_fn)
    fd = interior(dev.device_state().reshape(19, 16, 12, 14), 3)
    fo = interior(ref.f(), 3)
    err = rel_l2(fd, fo)
    ferr = max(rel_l2(a, b) for a, b in zip(fdv, fro))
    print(f"rel-L2(f)={err:.3e} rel-L2(F_atom)={ferr:.3e}")
    assert err <= TOL
    assert ferr <= 1e-11


def test_state_roundtrip_is_exact():
    import couplb_amd
    rng = np.random.default_rng(3)
    for dim, n in ((3, (9, 7, 5)), (2, (11, 6, 1))):
        d = couplb_amd.Device(dim, *n, tau=0.8, walls=((0, 0), (1, 2), (0, 0)))
        f = rng.uniform(0.01, 0.1, d.get_state().shape)
        d.set_state(f)
        g = d.get_state()
        fi, gi = interior(f, dim), interior(g, dim)
        assert np.array_equal(fi, gi)


@pytest.mark.parametrize("walls", [WALLS[0], WALLS[2], WALLS[4], WALLS[5]])
def test_fixed_markers_reuse_matches_oracle(walls):
    """Markers at fixed positions whose velocities change at every MD step.
    The fix passes positions_unchanged, so the device keeps the ordering, the
    coverage and the interpolated fluid velocity over the sub-steps of one LB
    interval and spreads the summed reaction once; the oracle recomputes
    everything at every sub-step as the reference does."""
    w = " ".join(f"wall_{ax} {walls[d][0]} {walls[d][1]}" for d, ax in enumerate("xyz"))
    cmd = (f"flow all couplb 12 10 14 0.1 1.0 {w} wall_vel 0.01 0.0 0.005 "
           f"gravity 1e-5 2e-6 -1e-6 kernel peskin4 md_per_lb 4 check_every 5")
    print("case fixed markers", cmd, flush=True)
    base = _ibm_atoms(60, 10.0, 11)

    def atoms_fn(step):
        a = copy_atoms(base)
        a["v"] = a["v"] * (1.0 + 0.5 * np.sin(0.7 * step))
        return a

    dev, fdv = run_device(cmd, 32, dimension=3, dt=0.25, atoms_fn=atoms_fn)
    ref, fro = run_oracle(cmd, 32, dimension=3, dt=0.25, atoms_fn=atoms_fn)
    fd = interior(dev.device_state().reshape(19, 16, 12, 14), 3)
    fo = interior(ref.f(), 3)
    err = rel_l2(fd, fo)
    ferr = max(rel_l2(a, b) for a, b in zip(fdv, fro))
    print(f"rel-L2(f)={err:.3e} rel-L2(F_atom)={ferr:.3e}")
    assert err <= TOL
    assert ferr <= 1e-11
    dd, od = dev.last_diag(), ref.last_diag()
    assert abs(dd["mass"] - od["mass"]) <= 1e-12 * od["mass"]


# reuse vs recompute bound per storage format: FP64 differs only in the
# spreading's summation order; a reduced format can also round a stored
# population to the neighbouring value where that order moved the sum
REUSE_TOL = {"fp64": 1e-14, "fp32": 1e-9, "fp16": 1e-5, "bf16": 1e-4}


@pytest.mark.parametrize("storage", ["fp64", "fp32", "fp16", "bf16"])
def test_device_reuse_matches_full_recompute(storage):
    """Device API: the same four sub-steps with positions declared unchanged
    (merged records, the fixed-marker map summed inside the step) and declared
    new (one record per sub-step, spread tile by tile) give the same
    populations and atom forces up to summation order."""
    import couplb_amd
    from tests.scenarios import DeviceArray
    rng = np.random.default_rng(5)
    n = 500
    x = rng.uniform(0.0, 24.0, (n, 3))
    res = []
    for reuse in (True, False):
        d = couplb_amd.Device(3, 24, 24, 24, tau=0.8, kernel=couplb_amd.PESKIN4, md_per_lb=4,
                              storage=storage)
        d.init_equilibrium(1.0, (0.02, 0.01, 0.0))
        dx, dm = DeviceArray(x), DeviceArray(np.ones(n))
        dv, df = DeviceArray(np.zeros((n, 3))), DeviceArray(np.zeros((n, 3)))
        kw = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0,
                  inv_N=0.25, force_scale=1.0)
        for step in range(3):
            d.step((1e-6, 0.0, 0.0), couplb_amd.RHO_REFRESHED_NEXT)
            for sub in range(4):
                d.sync()
                dv.upload(np.full((n, 3), 1e-3 * (sub + 1)))
                d.ibm_device(n, dx.addr, dv.addr, dm.addr, df.addr,
                             positions_unchanged=reuse and not (step == 0 and sub == 0), **kw)
        d.step((1e-6, 0.0, 0.0), couplb_amd.RHO_REFRESHED_NEXT)
        d.sync()
        res.append((d.get_state(), df.download()))
    (f1, a1), (f2, a2) = res
    print(f"{storage} reuse vs recompute: rel-L2(f) =", rel_l2(f1, f2), "rel-L2(F_atom) =", rel_l2(a1, a2))
    assert np.all(np.isfinite(f1)) and np.abs(a1).max() > 0
    assert rel_l2(f1, f2) <= REUSE_TOL[storage]
    assert rel_l2(a1, a2) <= REUSE_TOL[storage]


def test_group_growing_within_an_interval_matches_oracle():
    """The group gains atoms between the sub-steps of one LB interval
    (md_per_lb 2): the device grows its marker buffers while the earlier
    sub-step's records are still pending, and must spread them all, as the
    reference spreads every sub-step as it happens (fix_couplb.cpp:741-792)."""
    rng = np.random.default_rng(11)
    n_all = 240
    base = atoms_from(rng.uniform(0.0, 12.0, (n_all, 3)), rng.normal(0.0, 1e-3, (n_all, 3)),
                      rng.uniform(0.5, 1.5, n_all))
    sizes = {0: 50, 1: 50, 2: 51, 3: 51, 4: 110, 5: 60, 6: 240, 7: 240, 8: 30}

    def atoms(step):
        a = moving(base)(step)
        a["mask"] = (np.arange(n_all) < sizes[step]).astype(np.int32)
        return a

    cmd = "grow all couplb 12 12 12 0.1 1.0 md_per_lb 2 kernel peskin4 wall_y 1 2 wall_vel 0.01 0.0 0.0"
    print("case group growing within an interval", flush=True)
    dev, fd = run_device(cmd, 8, dimension=3, dt=0.5, atoms_fn=atoms)
    ref, fo = run_oracle(cmd, 8, dimension=3, dt=0.5, atoms_fn=atoms)
    a = interior(dev.device_state().reshape(19, 14, 14, 14), 3)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    ferr = max(rel_l2(p, q) for p, q in zip(fd, fo))
    print(f"rel-L2(f) = {err:.3e}, rel-L2(F_atom) = {ferr:.3e}")
    assert err <= TOL and ferr <= 1e-11


@pytest.mark.parametrize("md", [20, 37])
def test_more_sub_steps_than_record_slots_match_oracle(md):
    """md_per_lb above the 16 record slots (fix_couplb.cpp:111-114 accepts
    any md_per_lb >= 1): with moving markers every sub-step records, the
    full slots are spread into the node forces and reused."""
    rng = np.random.default_rng(md)
    n = 30
    base = atoms_from(rng.uniform(0.0, 12.0, (n, 3)), rng.normal(0.0, 1e-3, (n, 3)), rng.uniform(0.5, 1.5, n))

    def atoms(step):
        a = copy_atoms(base)
        a["x"] = a["x"] + 0.2 * np.sin(0.05 * step + np.arange(n))[:, None]
        return a

    cmd = f"many all couplb 12 12 12 0.1 1.0 md_per_lb {md} kernel roma wall_z 1 1 gravity 1e-5 0.0 0.0"
    print("case md_per_lb", md, flush=True)
    steps = 2 * md + 3
    dev, fd = run_device(cmd, steps, dimension=3, dt=1.0 / md, atoms_fn=atoms)
    ref, fo = run_oracle(cmd, steps, dimension=3, dt=1.0 / md, atoms_fn=atoms)
    a = interior(dev.device_state().reshape(19, 14, 14, 14), 3)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    ferr = max(rel_l2(p, q) for p, q in zip(fd, fo))
    print(f"rel-L2(f) = {err:.3e}, rel-L2(F_atom) = {ferr:.3e}")
    assert err <= TOL and ferr <= 1e-11


@pytest.mark.parametrize("reorder_every,stop", [("8", None), ("3", None), ("0", None), ("8", 45), (None, None),
                                                (None, 45)])
def test_moving_markers_across_cells_match_oracle(reorder_every, stop, monkeypatch):
    """Markers that cross cells and periodic faces at every few MD steps over
    40 LB intervals (md_per_lb 3): the device keeps a marker ordering for
    COUPLB_REORDER_EVERY intervals, sums the forces of the markers on their
    ordered base cells through the node-sorted map of that ordering, and
    spreads the records of markers that left their base cell tile by tile;
    the result must be the reference's (fix_couplb.cpp:723-798) to rounding.
    stop: the markers stand still from that MD step on (the kept ordering is
    renewed once it is old and the fixed-marker path takes over).  None: the
    default policy (renewed when old or when a quarter of the records left
    the ordering's base cells)."""
    if reorder_every is None:
        monkeypatch.delenv("COUPLB_REORDER_EVERY", raising=False)
    else:
        monkeypatch.setenv("COUPLB_REORDER_EVERY", reorder_every)
    rng = np.random.default_rng(31)
    n = 80
    nx, ny, nz = 20, 14, 12
    x0 = np.stack([rng.uniform(0, nx, n), rng.uniform(3, ny - 3, n), rng.uniform(0, nz, n)], 1)
    vx = rng.uniform(-0.15, 0.15, n)
    base = atoms_from(x0, rng.normal(0.0, 1e-3, (n, 3)), rng.uniform(0.5, 1.5, n))

    def atoms(step):
        a = copy_atoms(base)
        x = x0.copy()
        if stop is not None:
            step = min(step, stop)
        x[:, 0] = np.mod(x0[:, 0] + vx * step, nx)
        x[:, 1] = x0[:, 1] + 1.5 * np.sin(0.07 * step + np.arange(n))
        x[:, 2] = np.mod(x0[:, 2] + 0.6 * step * np.cos(np.arange(n)), nz)
        a["x"] = x
        return a

    cmd = (f"cross all couplb {nx} {ny} {nz} 0.1 1.0 md_per_lb 3 kernel peskin4 wall_y 1 2 "
           "wall_vel 0.01 0.0 0.0 gravity 1e-5 0.0 0.0")
    print("case moving markers across cells, COUPLB_REORDER_EVERY", reorder_every, "stop", stop, flush=True)
    steps = 120
    dev, fd = run_device(cmd, steps, dimension=3, dt=1.0 / 3, atoms_fn=atoms)
    ref, fo = run_oracle(cmd, steps, dimension=3, dt=1.0 / 3, atoms_fn=atoms)
    a = interior(dev.device_state().reshape(19, nz + 2, ny + 2, nx + 2), 3)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    ferr = max(rel_l2(p, q) for p, q in zip(fd, fo))
    print(f"rel-L2(f) = {err:.3e}, rel-L2(F_atom) = {ferr:.3e}")
    assert err <= TOL and ferr <= 1e-11


def test_host_array_flags_match_plain_host_calls():
    """couplb_ibm_substep with positions / masses declared unchanged and the
    forces declared zero (those arrays are then not sent) gives the same
    forces and the same state, bit for bit, as the calls that send every
    array, including a sub-step whose masses change."""
    import couplb_amd
    rng = np.random.default_rng(7)
    n, N, md = 40, 16, 2
    x0 = rng.uniform(2, N - 2, (n, 3))
    v = rng.normal(0.0, 1e-3, (n, 3))
    m = rng.uniform(0.5, 1.5, n)
    kw = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0, inv_N=1.0 / md, force_scale=1.0)
    devs = [couplb_amd.Device(3, N, N, N, tau=0.8, kernel=couplb_amd.PESKIN4, md_per_lb=md) for _ in range(2)]
    for d in devs:
        d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    prev = None
    for step in range(12):
        if step % md == 0:
            for d in devs:
                d.step((1e-6, 0.0, 0.0), couplb_amd.RHO_REFRESHED_NEXT)
        x = x0 + (0.3 * (step // 3)) * v / np.abs(v).max()   # positions change every third sub-step
        mm = m * (1.5 if step >= 7 else 1.0)                  # masses change once
        same_x = prev is not None and np.array_equal(prev, x)
        prev = x.copy()
        out = []
        for k, d in enumerate(devs):
            f = np.zeros((n, 3))
            if k == 0:
                same_m = step > 0 and step != 7
                d.ibm(x, v, mm, f, positions_unchanged=same_x, masses_unchanged=same_m, forces_zero=True, **kw)
            else:
                d.ibm(x, v, mm, f, positions_unchanged=same_x, **kw)
            out.append(f)
        assert np.array_equal(out[0], out[1]), step
    a, b = (d.get_state() for d in devs)
    assert np.array_equal(a, b)
    for d in devs:
        d.close()


def test_chunked_host_sub_steps_match_device_arrays():
    """Host-array sub-steps that only recompute forces (same positions, same
    LB state) send v (m, f where needed) in chunks and run each chunk as it
    arrives; forces and state equal the device-array path bit for bit, with
    forces returned as F or added to the caller's f and masses that change
    inside an interval."""
    import couplb_amd
    from tests.scenarios import DeviceArray
    rng = np.random.default_rng(11)
    n, N, md = 12000, 40, 4        # n >= 8192: four chunks
    kw = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0, inv_N=1.0 / md, force_scale=1.0)
    devs = [couplb_amd.Device(3, N, N, N, tau=0.8, kernel=couplb_amd.PESKIN4, md_per_lb=md) for _ in range(2)]
    for d in devs:
        d.init_equilibrium(1.0, (0.01, 0.0, 0.0))
    m = rng.uniform(0.5, 1.5, n)
    dx, dv, dm, df = (DeviceArray(np.zeros(sh)) for sh in ((n, 3), (n, 3), (n,), (n, 3)))
    for step in range(12):
        if step % md == 0:
            for d in devs:
                d.step((1e-6, 0.0, 0.0), couplb_amd.RHO_REFRESHED_NEXT)
            x = rng.uniform(2, N - 2, (n, 3))               # new positions every interval
        v = rng.normal(0.0, 1e-3, (n, 3))
        if step == 6:
            m = m * 1.25                                     # masses change inside an interval
        zero = step % 3 != 1
        f_in = np.zeros((n, 3)) if zero else rng.normal(0.0, 1e-4, (n, 3))
        unchanged = step % md != 0
        print("case chunked host sub-step", step, "forces_zero", zero, flush=True)
        fa = f_in.copy()
        devs[0].ibm(x, v, m, fa, positions_unchanged=unchanged, masses_unchanged=step not in (0, 6),
                    forces_zero=zero, **kw)
        for arr, val in ((dx, x), (dv, v), (dm, m), (df, f_in)):
            arr.upload(val)
        devs[1].ibm_device(n, dx.addr, dv.addr, dm.addr, df.addr, positions_unchanged=unchanged, **kw)
        devs[1].sync()
        assert np.array_equal(fa, df.download()), step
    a, b = (d.get_state() for d in devs)
    assert np.array_equal(a, b)
    for d in devs:
        d.close()
