"""Runs of LB steps on small grids: couplb_lbm_steps_async runs every step
after the first in one launch (k_steps_fused).  The fused run must leave the
bits that one launch per step leaves, whatever the first step brings in
(import form, a spread IBM force, a refreshed density for gravity), and the
fix, which holds back the steps of an empty group and issues them as runs,
must still give the oracle's results step for step."""
import numpy as np
import pytest

import couplb_amd
from tests.scenarios import atoms_from, copy_atoms, interior, rel_l2, run_device, run_oracle

pytestmark = pytest.mark.gpu

FACE = [[[0.0] * 3, [0.0] * 3] for _ in range(3)]
FACE[1][0] = [0.01, 0.0, 0.002]
FACE[1][1] = [-0.01, 0.0, 0.0]

# name: dim, extents, walls, face velocities, storage, gravity, steps
CASES = {
    "c1-d2q9": (2, (8, 32, 1), ((0, 0), (1, 1), (0, 0)), None, "fp64", (5e-6, 0, 0), 37),
    "d2q9-moving": (2, (16, 16, 1), ((0, 0), (2, 2), (0, 0)), FACE, "fp64", (0, 0, 0), 20),
    "d3q19-walls": (3, (8, 6, 6), ((3, 3), (2, 1), (0, 0)), FACE, "fp64", (1e-5, 0, 2e-6), 25),
    "d3q19-open": (3, (8, 8, 8), ((4, 4), (0, 0), (1, 1)), None, "fp64", (0, 2e-6, 0), 24),
    "d3q19-fp16": (3, (12, 10, 8), ((0, 0), (1, 1), (0, 0)), None, "fp16", (1e-5, 0, 0), 25),
    "d3q19-bf16": (3, (12, 10, 8), ((0, 0), (2, 2), (0, 0)), FACE, "bf16", (1e-5, 0, 0), 25),
    "d2q9-fp32": (2, (32, 32, 1), ((0, 0), (1, 1), (0, 0)), None, "fp32", (5e-6, 0, 0), 30),
}


def make(case):
    dim, n, walls, face, storage, g, steps = case
    d = couplb_amd.Device(dim, *n, tau=0.8, walls=walls, face_vel=face, storage=storage,
                          kernel=couplb_amd.PESKIN4)
    d.init_equilibrium(1.0, (0.01, -0.004, 0.003 if dim == 3 else 0.0))
    return d


def markers(n):
    rng = np.random.default_rng(3)
    x = rng.uniform([1.5, 2.5, 1.5], [n[0] - 1.5, n[1] - 2.5, max(n[2] - 1.5, 1.6)], (6, 3))
    v = rng.normal(0, 1e-3, x.shape)
    return x, v, np.ones(len(x))


@pytest.mark.parametrize("name", sorted(CASES))
def test_fused_run_equals_one_launch_per_step(name):
    case = CASES[name]
    dim, n, _, _, storage, g, steps = case
    print(f"case {name}: {dim}D {n} {storage} {steps} steps", flush=True)
    a, b = make(case), make(case)
    assert a.fuses_steps()
    # from the import form, then a run after a run, then after an IBM
    # sub-step (its force and refreshed density belong to the first step)
    for run, with_ibm in ((steps, False), (steps + 4, False), (11, True)):
        if with_ibm and dim == 3:
            x, v, m = markers(n)
            for d in (a, b):
                d.ibm(x, v, m, np.zeros((len(m), 3)), (0, 0, 0), 1.0, 1.0, 1.0, 1.0, 1.0)
        a.steps_async(run, g)
        for _ in range(run):
            b.steps_async(1, g)
        fa, fb = a.get_state(), b.get_state()
        print(f"  run of {run}{' after IBM' if with_ibm else ''}: identical {np.array_equal(fa, fb)}",
              flush=True)
        assert np.array_equal(fa, fb)
    assert np.isfinite(fa).all()


def test_large_grids_run_step_by_step():
    # more nodes than the bound, or two states beyond the LDS of a workgroup
    assert not couplb_amd.Device(2, 64, 64, 1, tau=0.8).fuses_steps()
    assert not couplb_amd.Device(3, 12, 10, 8, tau=0.8).fuses_steps()
    assert couplb_amd.Device(3, 12, 10, 8, tau=0.8, storage="fp16").fuses_steps()


def test_fix_held_back_steps_match_oracle_as_atoms_join():
    # group empty for the first 23 steps (steps held back, issued at the
    # check steps, at the joining atoms' first IBM sub-step and at the end)
    cmd = ("flow all couplb 8 6 6 0.1 1.0 wall_y 1 1 gravity 1e-5 2e-6 0 "
           "kernel peskin4 md_per_lb 2 check_every 7")
    rng = np.random.default_rng(11)
    x0 = rng.uniform([1.5, 1.5, 1.5], [6.5, 4.5, 4.5], (5, 3))
    v0 = rng.normal(0, 1e-3, x0.shape)
    base = atoms_from(x0, v0, np.ones(5))
    print("case fix 8x6x6: atoms join at step 24, 40 steps", flush=True)

    def atoms_fn(step):
        a = copy_atoms(base)
        a["mask"][:] = 1 if step >= 24 else 0
        return a

    dev, fdev = run_device(cmd, 40, dt=0.5, atoms_fn=atoms_fn)
    assert couplb_amd.lib().couplb_lbm_fuses_steps(couplb_amd.lib().couplb_fix_context(dev.h)) == 1
    ref, fref = run_oracle(cmd, 40, dt=0.5, atoms_fn=atoms_fn)
    fd = interior(dev.device_state().reshape(19, 8, 8, 10), 3)
    fo = interior(ref.f(), 3)
    err = rel_l2(fd, fo)
    print("  rel-L2(f) =", err, flush=True)
    assert err <= 1e-12
    assert np.allclose(np.array(fdev), np.array(fref), rtol=1e-10, atol=1e-16)
    dd, od = dev.last_diag(), ref.last_diag()
    assert abs(dd["mass"] - od["mass"]) <= 1e-12 * od["mass"]


def test_fix_held_back_steps_with_variable_gravity():
    # an equal-style gravity variable changes every step: each held-back run
    # is one step long, and the results are the oracle's bit for bit
    cmd = "flow all couplb 8 32 1 0.1 1.0 wall_y 1 1 gravity v_gx 0.0 0.0 check_every 50"
    gx = lambda step: 5e-6 * (1.0 + 0.01 * (step % 17))
    import oracle
    from tests.scenarios import parse_cmd
    fd = couplb_amd.Fix(cmd, dimension=2, dt=1.0, variables={"gx": gx})
    fd.init()
    kw = parse_cmd(cmd.replace("v_gx", "0.0"))
    fo = oracle.Fix(dim=2, dt=1.0, **kw)
    fd.setup(None)
    fo.setup(None)
    print("case C1 deck with gravity v_gx, 120 steps", flush=True)
    for step in range(1, 121):
        fd.step = fo.step = step
        fo.set_gravity((gx(step), 0.0, 0.0))
        fd.post_force(None)
        fo.post_force(None)
        fd.end_of_step(step)
        fo.end_of_step(step)
    a = interior(fd.device_state().reshape(9, 1, 34, 10), 2)
    b = interior(fo.f(), 2)
    assert np.array_equal(a, b)
    fd.close()
    fo.close()
