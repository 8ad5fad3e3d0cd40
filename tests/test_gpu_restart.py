"""Checkpoint and restart through the fix (couplb_io.h:653-828): the device
writes the file in the reference's layout, which the oracle's own reader
accepts and which holds the device state; a fix restarted from it continues
as the oracle fix continues from the same populations (restart = the state
set and the macroscopic fields recomputed, couplb_io.h:817, then setup)."""
import numpy as np
import pytest

import couplb_amd
import oracle
from tests.scenarios import atoms_from, interior, parse_cmd, rel_l2

pytestmark = pytest.mark.gpu

N = (12, 10, 8)
SHAPE = (19, N[2] + 2, N[1] + 2, N[0] + 2)   # reference layout (q, z, y, x), ghosts included
BASE = ("flow all couplb 12 10 8 0.1 1.0 wall_y 1 2 wall_vel 0.01 0.0 0.0 gravity 1e-5 2e-6 0.0 "
        "kernel peskin4 md_per_lb 2")


def markers():
    rng = np.random.default_rng(21)
    x = rng.uniform([1.0, 2.5, 1.0], [11.0, 7.5, 7.0], (7, 3))
    v = rng.normal(0, 1e-3, x.shape)
    return atoms_from(x, v, np.ones(len(x)))


def drive(fx, a, first, last, forces=None):
    for step in range(first, last + 1):
        fx.step = step
        a["f"][:] = 0.0
        fx.post_force(a)
        fx.end_of_step(step)
        if forces is not None:
            forces.append(a["f"].copy())


def test_checkpoint_file_and_restart_match_oracle(tmp_path):
    pre = str(tmp_path / "ck")
    print("case restart: 12x10x8, moving wall, gravity, 7 markers; checkpoint at step 10", flush=True)
    a = markers()
    fd = couplb_amd.Fix(BASE + f" checkpoint 10 {pre}", dt=0.5)
    fd.init()
    fd.setup(a)
    drive(fd, a, 1, 10)
    state = fd.device_state().reshape(SHAPE)
    fd.close()
    step, f = oracle.read_checkpoint(pre + ".0.clbk", 3, 19, N, N, state.size)
    assert step == 10
    f = f.reshape(state.shape)
    assert np.array_equal(interior(f, 3), interior(state, 3))

    b_dev, b_ora = markers(), markers()
    fr = couplb_amd.Fix(BASE + f" restart {pre}", dt=0.5, ntimestep=10)
    fr.init()
    fo = oracle.Fix(dim=3, dt=0.5, ntimestep=10, **parse_cmd(BASE))
    fo.set_f(f)
    fr.setup(b_dev)
    fo.setup(b_ora)
    fdv, fov = [], []
    drive(fr, b_dev, 11, 20, fdv)
    drive(fo, b_ora, 11, 20, fov)
    a1 = interior(fr.device_state().reshape(SHAPE), 3)
    a2 = interior(fo.f(), 3)
    err = rel_l2(a1, a2)
    ferr = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(fdv, fov))
    print(f"  after restart: rel-L2(f) = {err:.3e}, per-atom forces {ferr:.3e}", flush=True)
    assert err <= 1e-12
    assert ferr <= 1e-11
    fr.close()
    fo.close()


@pytest.mark.parametrize("with_markers", [False, True])
def test_checkpoint_file_equals_the_oracle_file_ghosts_included(tmp_path, with_markers):
    # The reference writes the whole ghost-inclusive f (couplb_io.h:706): its
    # periodic ghost planes hold the images of the post-collision populations
    # of the step before the last one (exchange before the stream, then the
    # swap, couplb_streaming.h:133-182), its wall ghost planes the
    # init_equilibrium values.  Without markers the whole file is identical;
    # with them the populations differ by the spreading's summation order.
    pre = str(tmp_path / "ck")
    cmd = BASE + f" checkpoint 10 {pre}"
    print("case checkpoint file vs oracle file, markers", with_markers, flush=True)
    a_dev = markers() if with_markers else None
    a_ora = markers() if with_markers else None
    fd = couplb_amd.Fix(cmd, dt=0.5)
    fd.init()
    fd.setup(a_dev)
    fo = oracle.Fix(dim=3, dt=0.5, **parse_cmd(BASE))
    fo.setup(a_ora)
    for step in range(1, 11):
        for fx, a in ((fd, a_dev), (fo, a_ora)):
            fx.step = step
            if a is not None:
                a["f"][:] = 0.0
            fx.post_force(a)
            fx.end_of_step(step)
    fd.close()
    ref = str(tmp_path / "oracle.clbk")
    assert oracle.write_checkpoint(ref, 3, 19, N, N, 10, fo.f())
    got, want = open(pre + ".0.clbk", "rb").read(), open(ref, "rb").read()
    hdr = len(want) - 8 * int(np.prod(SHAPE))
    assert len(got) == len(want) and got[:hdr] == want[:hdr]      # header
    g = np.frombuffer(got[hdr:], np.float64).reshape(SHAPE)
    w = np.frombuffer(want[hdr:], np.float64).reshape(SHAPE)
    ghost = np.ones(SHAPE[1:], bool)
    ghost[1:-1, 1:-1, 1:-1] = False
    print("  ghost planes rel-L2", rel_l2(g[:, ghost], w[:, ghost]), "bit-exact file", got == want)
    if with_markers:
        assert rel_l2(g, w) <= 1e-12 and rel_l2(g[:, ghost], w[:, ghost]) <= 1e-12
    else:
        assert got == want
    fo.close()
