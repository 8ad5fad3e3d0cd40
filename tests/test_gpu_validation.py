"""The reference's own validation decks, run through the MI355X fix with a
velocity-Verlet integrator in place of LAMMPS' fix nve (per step:
initial_integrate, periodic remap at the neighbour-list rebuilds, f = 0,
post_force of fix addforce and of the fix, final_integrate, end_of_step):

- tests/subcycling/validate.sh:320-390: with md_per_lb 4 and 10 the particle
  velocity tracks the md_per_lb 1 run within 2 %, v_x > 0 at the end,
  max |v_y| < 0.01, the first printed |f_x| lies in (0.1, 20); the first
  steps of the md_per_lb 1 run also match the oracle fix driven the same way;
- tests/drag_forces_point/in.drag_point:1-15, 89-93: no transverse motion,
  constant mass, Ma < 0.1, no density clamps, v_x > 0, and fluid + particle
  momentum equal to the applied impulse (with the half-step terms of the
  integrator);
- tests/drag_forces_sphere/in.drag_sphere:46-50: terminal velocity of the
  162-marker sphere (fix rigid/nve single: translation only here) against
  Hasimoto's drag for a periodic array of spheres;
- tests/poiseuille3d/in.poiseuille3d with validate.py's rule on the written
  profile file.
"""
import os

import numpy as np
import pytest

import couplb_amd
import oracle
from tests.scenarios import parse_cmd

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def verlet(fix, atoms, nsteps, dt, fext=None, every=1, lo=None, hi=None, periodic=(1, 1, 1),
           rigid=False, on_step=None):
    """LAMMPS Verlet order for the atoms of `atoms` (all in the group).
    rigid: one body (fix rigid/nve single) moving without rotation."""
    m = atoms["mass"][:, None]
    M = atoms["mass"].sum()
    for s in range(1, nsteps + 1):
        fix.step += 1
        if rigid:
            V = atoms["v"][0] + 0.5 * dt * atoms["f"].sum(0) / M
            atoms["v"][:] = V
            atoms["x"] += dt * V
        else:
            atoms["v"] += 0.5 * dt * atoms["f"] / m
            atoms["x"] += dt * atoms["v"]
        if s % every == 0 and lo is not None:
            for d in range(3):
                if periodic[d]:
                    L = hi[d] - lo[d]
                    atoms["x"][:, d] = lo[d] + np.mod(atoms["x"][:, d] - lo[d], L)
        atoms["f"][:] = 0.0
        if fext is not None:
            atoms["f"] += fext
        fix.post_force(atoms)
        if rigid:
            atoms["v"][:] = atoms["v"][0] + 0.5 * dt * atoms["f"].sum(0) / M
        else:
            atoms["v"] += 0.5 * dt * atoms["f"] / m
        fix.end_of_step(fix.step)
        if on_step is not None:
            on_step(s)


def one_atom(x, v, m):
    return dict(x=np.array([x], float), v=np.array([v], float), f=np.zeros((1, 3)),
                mask=np.ones(1, np.int32), mass=np.array([m], float))


# tests/subcycling/validate.sh:22-45
DX, NX, NY, NU, RHO, G0, DT_LB, PMASS, V0X = 0.625, 96, 64, 0.1, 1.26, 2e-4, 0.08333, 0.161, -0.3
LX, LY = NX * DX, NY * DX
SUB_BOX = dict(dimension=2, boxlo=(0.0, 0.0, -0.5), boxhi=(LX, LY, 0.5), periodicity=(1, 0, 1))


def sub_deck(md, extra):
    md_kw = f" md_per_lb {md}" if md > 1 else ""
    return (f"flow all couplb {NX} {NY} 1 {NU} {RHO}{md_kw} wall_y 1 1 gravity {G0} 0.0 0.0 "
            f"kernel roma {extra}")


def track(md, ckpt, steps_lb=2000, print_lb=10):
    """validate.sh phase 2: the particle from the domain centre at v0x,
    printed at setup and every print_lb LB steps as (t, x, y, vx, vy, fx, fy)."""
    dt = DT_LB / md
    fx = couplb_amd.Fix(sub_deck(md, f"restart {ckpt} check_every {steps_lb * md // 2}"), dt=dt,
                        **SUB_BOX)
    fx.init()
    a = one_atom((LX / 2, LY / 2, 0.0), (V0X, 0.0, 0.0), PMASS)
    fx.setup(a)
    # fix print writes at setup (step 0) and every print_lb * md steps
    rows = [(0.0, a["x"][0, 0], a["x"][0, 1], a["v"][0, 0], a["v"][0, 1], a["f"][0, 0], a["f"][0, 1])]

    def rec(s):
        if s % (print_lb * md) == 0:
            rows.append((s * dt, a["x"][0, 0], a["x"][0, 1], a["v"][0, 0], a["v"][0, 1],
                         a["f"][0, 0], a["f"][0, 1]))

    verlet(fx, a, steps_lb * md, dt, every=md, lo=(0.0, 0.0, -0.5), hi=(LX, LY, 0.5),
           periodic=(1, 0, 1), on_step=rec)
    fx.close()
    return np.array(rows)


def test_subcycling_matches_baseline(tmp_path):
    ckpt = str(tmp_path / "ckpt_pois")
    print("case validate.sh spin-up: D2Q9 96x64 Poiseuille, 10000 steps", flush=True)
    fx = couplb_amd.Fix(sub_deck(1, f"check_every 5000 checkpoint 10000 {ckpt}"), dt=DT_LB, **SUB_BOX)
    fx.init()
    fx.setup(None)
    for _ in range(10000):
        fx.step += 1
        fx.post_force(None)
        fx.end_of_step(fx.step)
    fx.close()
    assert os.path.exists(ckpt + ".0.clbk")
    runs = {}
    for md in (1, 4, 10):
        print(f"case validate.sh track md_per_lb {md}", flush=True)
        runs[md] = track(md, ckpt)
    base = runs[1]
    assert len(base) == 201 and all(len(r) == 201 for r in runs.values())
    # TEST 1 (validate.sh:306-318): the particle ends moving rightward
    assert base[-1, 3] > 0
    # TEST 2 (validate.sh:320-377): max |dvx / vx| over points 6.. with |vx| >= 0.01
    for md in (4, 10):
        vb, vt = base[:, 3], runs[md][:, 3]
        keep = np.arange(len(vb)) >= 5
        keep &= np.abs(vb) >= 0.01
        err = np.max(np.abs(vt[keep] - vb[keep]) / np.abs(vb[keep]))
        print(f"  md_per_lb {md}: max |dvx/vx| = {err:.3e}, final {vt[-1]:.6f} vs {vb[-1]:.6f}",
              flush=True)
        assert err < 0.02
    # TEST 3 (validate.sh:379-388): no vertical drift
    assert np.max(np.abs(base[:, 4])) < 0.01
    # TEST 4 (validate.sh:391-400): initial force magnitude O(1)
    assert 0.1 < abs(base[0, 5]) < 20


def test_subcycling_deck_matches_oracle_from_rest():
    # the md_per_lb 4 deck from a fluid at rest, 400 MD steps, device vs oracle
    md, dt, n = 4, DT_LB / 4, 400
    cmd = sub_deck(md, "check_every 100")
    print("case validate.sh deck md_per_lb 4 from rest: device vs oracle", flush=True)
    fd = couplb_amd.Fix(cmd, dt=dt, **SUB_BOX)
    fd.init()
    kw = parse_cmd(cmd)
    fo = oracle.Fix(dim=2, dt=dt, boxlo=(0.0, 0.0, -0.5), boxhi=(LX, LY, 0.5), **kw)
    traj = {}
    for name, fx in (("dev", fd), ("ora", fo)):
        a = one_atom((LX / 2, LY / 2, 0.0), (V0X, 0.0, 0.0), PMASS)
        fx.setup(a)
        rows = []
        verlet(fx, a, n, dt, every=md, lo=(0.0, 0.0, -0.5), hi=(LX, LY, 0.5), periodic=(1, 0, 1),
               on_step=lambda s: rows.append(np.concatenate([a["x"][0], a["v"][0], a["f"][0]])))
        traj[name] = np.array(rows)
    d, o = traj["dev"], traj["ora"]
    err = np.max(np.abs(d - o) / (np.abs(o).max(axis=0) + 1e-300))
    print(f"  max relative trajectory difference {err:.3e}", flush=True)
    assert err < 1e-10
    fd.close()
    fo.close()


def test_point_force_drag_invariants():
    # tests/drag_forces_point/in.drag_point: 32^3, roma, F = 1e-3 along x,
    # nu = 0.1, dt = 1 (lattice units), 10000 steps
    N, F, dt, steps = 32, 1e-3, 1.0, 10000
    cmd = f"flow all couplb {N} {N} {N} 0.1 1.0 kernel roma check_every 500"
    print("case in.drag_point: 32^3, 10000 steps", flush=True)
    fx = couplb_amd.Fix(cmd, dt=dt, boxlo=(-16.0, -16.0, -16.0), boxhi=(16.0, 16.0, 16.0))
    fx.init()
    a = one_atom((0.0, 0.0, 0.0), (0.0, 0.0, 0.0), 1.0)
    fext = np.array([[F, 0.0, 0.0]])
    a["f"][:] = fext
    fx.setup(a)
    f_ibm = [a["f"][0, 0] - F]          # setup's sub-step
    diags = []
    mass0 = float(N ** 3)

    def rec(s):
        f_ibm.append(a["f"][0, 0] - F)
        if s % 500 == 0:
            d = fx.last_diag()
            # fluid momentum: reactions of sub-steps 0..s-1; particle: Verlet
            # sums with half weights at both ends (m v_0 = 0)
            pred = F * s * dt + 0.5 * dt * (f_ibm[s] - f_ibm[0])
            diags.append((s, d["Ma"], d["mass"], d["mom"][0], d["clamps"], a["v"][0].copy(), pred))

    verlet(fx, a, steps, dt, fext=fext, lo=(-16.0,) * 3, hi=(16.0,) * 3, on_step=rec)
    fx.close()
    assert len(diags) == steps // 500
    for s, ma, mass, mom, clamps, v, pred in diags:
        assert ma < 0.1
        assert clamps == 0
        assert abs(mass - mass0) <= 1e-11 * mass0
        assert abs(v[1]) < 1e-8 and abs(v[2]) < 1e-8
        total = mom + 1.0 * v[0]
        assert abs(total - pred) <= 1e-9 * abs(pred), (s, total, pred)
    s, _, _, _, _, v, _ = diags[-1]
    print(f"  v_x at step {s}: {v[0]:.6e}; total momentum {diags[-1][3] + v[0]:.6e} "
          f"(impulse {F * s:.6e})", flush=True)
    assert v[0] > 0


def test_sphere_drag_hasimoto():
    # tests/drag_forces_sphere/in.drag_sphere: R = 3, 162 markers, 96^3 cells
    # of dx = 0.5 (L = 48), M = 50, F = 0.005 along x with a compensating
    # body force, dt = 0.25, roma, 30000 steps
    Nc, R, L, F, dt, steps = 96, 3.0, 48.0, 0.005, 0.25, 30000
    xs = np.load(os.path.join(GOLD, "sphere_markers.npz"))["x"]
    n = len(xs)
    g_comp = -F / (1.0 * L ** 3)
    cmd = f"flow all couplb {Nc} {Nc} {Nc} 0.1 1.0 kernel roma gravity v_g_comp 0.0 0.0 check_every 2000"
    print("case in.drag_sphere: 96^3, 162 markers, 30000 steps", flush=True)
    fx = couplb_amd.Fix(cmd, dt=dt, boxlo=(-24.0,) * 3, boxhi=(24.0,) * 3, variables={"g_comp": g_comp})
    fx.init()
    a = dict(x=xs.copy(), v=np.zeros((n, 3)), f=np.zeros((n, 3)), mask=np.ones(n, np.int32),
             mass=np.full(n, 50.0 / n))
    fext = np.zeros((n, 3))
    fext[:, 0] = F / n
    a["f"][:] = fext
    fx.setup(a)
    vx = []
    verlet(fx, a, steps, dt, fext=fext, lo=(-24.0,) * 3, hi=(24.0,) * 3, rigid=True,
           on_step=lambda s: vx.append(a["v"][0, 0]) if s % 1000 == 0 else None)
    d = fx.last_diag()
    fx.close()
    a_over_L = R / L
    K = 1.0 - 2.837 * a_over_L + 4.19 * a_over_L ** 3
    U = F / (6.0 * np.pi * 0.1 * R) * K
    err = (vx[-1] - U) / U
    print(f"  U_measured {vx[-1]:.6e}  U_Hasimoto {U:.6e}  error {100 * err:.2f} %  "
          f"(last 5000 steps: {vx[-6]:.6e} -> {vx[-1]:.6e}); Ma {d['Ma']:.3e}", flush=True)
    assert d["clamps"] == 0
    assert abs(vx[-1] - vx[-6]) <= 0.01 * abs(vx[-1])   # terminal
    assert abs(err) < 0.10


def read_profile(path):
    """validate.py:29-72 restated: header parameters and the last block."""
    import re
    params, steps, cur = {}, {}, None
    with open(path) as fh:
        lines = fh.read().splitlines()
    for s in (ln.strip() for ln in lines):
        if not s.startswith("#"):
            break
        m = re.search(r"H=Ny=(\d+)", s)
        if m:
            params["Ny"] = int(m.group(1))
        for key in ("tau", "nu", "nu_lb", "rho"):
            m = re.search(key + r"=([\d.eE+\-]+)", s)
            if m:
                params[key] = float(m.group(1))
    for s in (ln.strip() for ln in lines):
        if not s:
            continue
        if s.startswith("# step ="):
            cur = int(s.split("=")[1])
            steps[cur] = []
            continue
        if s.startswith("#") or cur is None:
            continue
        parts = s.split()
        if len(parts) >= 6:
            steps[cur].append([float(x) for x in parts])
    last = max(steps)
    return params, last, np.array(steps[last])


def test_poiseuille3d_deck(tmp_path):
    # tests/poiseuille3d/in.poiseuille3d: D3Q19 8x32x8, walls in y, g_x = 5e-6,
    # 50000 steps, profile every 5000; validate.py's rule: max |u - u_exact| <
    # 1 % of u_max.  The first 2000 steps also match the oracle bit for bit.
    out = tmp_path / "couplb_poiseuille3d.dat"
    cmd = (f"flow noatoms couplb 8 32 8 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 "
           f"output 5000 {out} check_every 2000")
    box = dict(boxlo=(0.0, 0.0, 0.0), boxhi=(8.0, 32.0, 8.0), periodicity=(1, 0, 1))
    print("case in.poiseuille3d: D3Q19 8x32x8, 50000 steps", flush=True)
    fd = couplb_amd.Fix(cmd, dt=1.0, **box)
    fd.init()
    fd.setup(None)
    kw = parse_cmd(cmd)
    kw.pop("output_every", None)
    fo = oracle.Fix(dim=3, dt=1.0, boxlo=box["boxlo"], boxhi=box["boxhi"], **kw)
    fo.setup(None)
    for step in range(1, 50001):
        fd.step = step
        fd.post_force(None)
        fd.end_of_step(step)
        if step <= 2000:
            fo.step = step
            fo.post_force(None)
            fo.end_of_step(step)
        if step == 2000:
            a = fd.device_state().reshape(19, 10, 34, 10)[:, 1:-1, 1:-1, 1:-1]
            b = fo.f().reshape(19, 10, 34, 10)[:, 1:-1, 1:-1, 1:-1]
            assert np.array_equal(a, b)
            md, mo = fd.last_diag()["mass"], fo.last_diag()["mass"]
            assert abs(md - mo) <= 1e-12 * mo
            fo.close()
    fd.close()
    params, last, data = read_profile(out)
    assert params["Ny"] == 32 and last == 50000 and len(data) == 32
    y, ux = data[:, 2], data[:, 4]
    dx = y[1] - y[0]
    H = params["Ny"] * dx
    exact = 5e-6 / (2 * params["nu"]) * (y + 0.5 * dx) * (H - 0.5 * dx - y)
    rel = np.max(np.abs(ux - exact)) / exact.max()
    print(f"  step {last}: max |u - u_exact| / u_max = {100 * rel:.4f} %", flush=True)
    assert rel < 0.01


def test_poiseuille2d_deck(tmp_path):
    # tests/poiseuille2d/in.poiseuille2d:21-30: D2Q9 16x64, walls in y, g_x =
    # 5e-6, Roma kernel, an empty group, 20000 steps, profile every 400,
    # check every 200 (box 16 x 64, z from -0.5 to 0.5, boundary p f p).
    # The whole profile file equals the oracle fix's, and validate.py's rule
    # (validate.py:75-111) passes: max |u - u_exact| < 1 % of u_max.
    out, ref = tmp_path / "couplb_poiseuille.dat", tmp_path / "oracle_poiseuille.dat"
    deck = "flow noatoms couplb 16 64 1 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 kernel roma "
    box = dict(boxlo=(0.0, 0.0, -0.5), boxhi=(16.0, 64.0, 0.5), periodicity=(1, 0, 1))
    print("case in.poiseuille2d: D2Q9 16x64, 20000 steps", flush=True)
    fd = couplb_amd.Fix(deck + f"output 400 {out} check_every 200", dimension=2, dt=1.0, **box)
    fd.init()
    fd.setup(None)
    kw = parse_cmd(deck + "check_every 200")
    fo = oracle.Fix(dim=2, dt=1.0, boxlo=box["boxlo"], boxhi=box["boxhi"], output_every=400,
                    output_file=str(ref), **kw)
    fo.setup(None)
    for step in range(1, 20001):
        for fx in (fd, fo):
            fx.step = step
            fx.post_force(None)
            fx.end_of_step(step)
    a = fd.device_state().reshape(9, 1, 66, 18)[:, :, 1:-1, 1:-1]
    b = fo.f()[:, :, 1:-1, 1:-1]
    assert np.array_equal(a, b)
    fd.close()
    fo.close()
    assert out.read_text() == ref.read_text()
    params, last, data = read_profile(out)
    assert params["Ny"] == 64 and last == 20000 and len(data) == 64
    y, ux = data[:, 2], data[:, 4]
    dx = y[1] - y[0]
    H = params["Ny"] * dx
    exact = 5e-6 / (2 * params["nu"]) * (y + 0.5 * dx) * (H - 0.5 * dx - y)
    rel = np.max(np.abs(ux - exact)) / exact.max()
    print(f"  step {last}: max |u - u_exact| / u_max = {100 * rel:.4f} % (u_max {exact.max():.4e})",
          flush=True)
    assert rel < 0.01
