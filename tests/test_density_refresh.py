"""Which density the gravity force of an LB step multiplies.

apply_external_force adds g * rho[n] with whatever the density array holds
(fix_couplb.cpp:680-708).  That array is written by the step's own
compute_macroscopic(true) (fix_couplb.cpp:663, after the force), by
compute_macroscopic(false) at the first IBM sub-step with group atoms
(757-760), at check / output / VTK steps of end_of_step (859, 880) and by a
restart (couplb_io.h:817).  So in a run without atoms the force of LB step m
uses the density of step m-1's pre-collision state, unless an end_of_step
after step m-1 ran a check, which stores the density of the current state.

The force is read back from the populations alone: in a periodic box the
post-collision populations are the post-stream ones shifted back along e_q
(an exact permutation), and BGK with Guo forcing changes the first moment of
every node by exactly F (couplb_collision.h:36-58: -omega (j - rho u) =
omega F / 2 with u = (j + F/2)/rho, and sum_q e_q S_q = (1 - omega/2) F).
So F(x) / g is the density the step used, and it is compared with the two
candidates.  The CPU test pins the oracle's orchestration; the GPU test runs
the same trace through the MI355X fix.
"""
import numpy as np
import pytest

import oracle
from tests.scenarios import interior

G = (2e-4, -1e-4, 5e-5)


def density_wave(dim, n):
    """Ghost-inclusive equilibrium populations of a density and velocity wave."""
    nx, ny, nz = n
    k = np.arange(nz if dim == 3 else 1)[:, None, None]
    j = np.arange(ny)[None, :, None]
    i = np.arange(nx)[None, None, :]
    rho = 1.0 + 0.05 * np.sin(2 * np.pi * i / nx) + 0.03 * np.cos(2 * np.pi * j / ny) + 0.0 * k
    u = np.zeros((3,) + rho.shape)
    u[0] = 0.01 * np.sin(2 * np.pi * j / ny) + 0.0 * i + 0.0 * k
    u[1] = 0.005 * np.cos(2 * np.pi * i / nx) + 0.0 * j + 0.0 * k
    if dim == 3:
        u[2] = 0.004 * np.sin(2 * np.pi * k / nz) + 0.0 * i + 0.0 * j
    f = oracle.feq(dim, rho, u)
    if dim == 3:
        return np.pad(f, ((0, 0), (1, 1), (1, 1), (1, 1)), mode="wrap")
    return np.pad(f, ((0, 0), (0, 0), (1, 1), (1, 1)), mode="wrap")


def guo_force(before, after, dim):
    """Node force of one LB step in a periodic box, [3, nz, ny, nx]."""
    e, _ = oracle.lattice(dim)
    a, b = interior(after, dim), interior(before, dim)
    F = np.zeros((3,) + a.shape[1:])
    for q in range(len(e)):
        # post-collision f*_q(x) = post-stream f_q(x + e_q)
        s = np.roll(a[q], shift=(-e[q][2], -e[q][1], -e[q][0]), axis=(0, 1, 2))
        d = s - b[q]
        for c in range(3):
            if e[q][c]:
                F[c] += e[q][c] * d
    return F


def density(f, dim):
    return interior(f, dim).sum(axis=0)


def refresh_trace(fix, state, steps, md_per_lb, check_every):
    """Drive an initialised fix without atoms through `steps` MD steps in
    LAMMPS' order and return, per LB step, (F, expected density, the other
    candidate or None, whether the density was refreshed)."""
    out = []
    fresh = True                  # the import stored the current density
    prev = None
    next_check = check_every if check_every > 0 else None
    for s in range(1, steps + 1):
        if (s - 1) % md_per_lb == 0:
            before = state()
            fix.post_force(None)
            after = state()
            now = density(before, fix_dim(fix))
            exp, other = (now, prev) if fresh else (prev, now)
            out.append((guo_force(before, after, fix_dim(fix)), exp, other, fresh))
            prev = now
            fresh = False
        else:
            fix.post_force(None)
        fix.end_of_step(s)
        if next_check is not None and s >= next_check:
            fresh = True
            next_check = s + check_every
    return out


def fix_dim(fix):
    return getattr(fix, "dim", None) or getattr(fix, "dimension")


def check_trace(trace, dim):
    g = np.array(G)
    tol = 1e-12 * np.abs(g).max()
    discriminated = 0
    for m, (F, exp, other, _) in enumerate(trace, start=1):
        for c in range(dim):
            err = np.abs(F[c] - g[c] * exp).max()
            assert err < tol, (m, c, err)
        if other is not None:
            gap = np.abs(F[0] - g[0] * other).max()
            # the other candidate would miss by far more than rounding
            assert gap > 1e4 * tol, (m, gap)
            discriminated += 1
    assert discriminated >= len(trace) - 1


# (dim, grid, md_per_lb, check_every); the MD timestep is dt_LB / md_per_lb
# with dt_LB = dx = 1, so the lattice gravity is G
CASES = [
    (3, (10, 8, 6), 1, 0),
    (3, (10, 8, 6), 1, 3),
    (3, (10, 8, 6), 2, 3),
    (2, (16, 12, 1), 1, 4),
]


@pytest.mark.parametrize("dim,n,md,chk", CASES)
def test_oracle_gravity_reads_the_density_the_reference_holds(dim, n, md, chk):
    fx = oracle.Fix(*n, 0.1, 1.0, dim=dim, md_per_lb=md, gravity=G, check_every=chk, dt=1.0 / md)
    fx.set_f(density_wave(dim, n))
    fx.setup(None)
    trace = refresh_trace(fx, fx.f, 8 * md, md, chk)
    check_trace(trace, dim)
    # with checks both branches ran; without, only the import's refresh
    refreshed = sum(1 for t in trace[1:] if t[3])
    assert (refreshed > 0 and refreshed < len(trace) - 1) if chk else refreshed == 0


def test_a_step_reading_the_current_density_fails_the_check():
    # the check tells the two candidates apart: feeding it the fresh density
    # where the stale one is due makes it fail
    dim, n = 3, (10, 8, 6)
    fx = oracle.Fix(*n, 0.1, 1.0, dim=dim, gravity=G)
    fx.set_f(density_wave(dim, n))
    fx.setup(None)
    trace = refresh_trace(fx, fx.f, 4, 1, 0)
    swapped = [trace[0]] + [(F, other, exp, fr) for F, exp, other, fr in trace[1:]]
    with pytest.raises(AssertionError):
        check_trace(swapped, dim)


@pytest.mark.gpu
@pytest.mark.parametrize("dim,n,md,chk", CASES + [(3, (40, 8, 6), 1, 3)])
def test_device_gravity_reads_the_density_the_reference_holds(dim, n, md, chk):
    import couplb_amd
    cmd = (f"rf all couplb {n[0]} {n[1]} {n[2]} 0.1 1.0 md_per_lb {md} "
           f"gravity {G[0]} {G[1]} {G[2]}" + (f" check_every {chk}" if chk else ""))
    print("case density refresh", dim, n, md, chk, flush=True)
    fx = couplb_amd.Fix(cmd, dimension=dim, dt=1.0 / md)
    fx.init()
    gd = (n[0] + 2, n[1] + 2, n[2] + 2 if dim == 3 else 1)
    Q = 19 if dim == 3 else 9
    import ctypes as C
    ctx = couplb_amd.lib().couplb_fix_context(fx.h)
    f0 = np.ascontiguousarray(density_wave(dim, n).reshape(-1))
    assert couplb_amd.lib().couplb_set_state_f(ctx, f0.ctypes.data_as(C.POINTER(C.c_double))) == 0
    fx.setup(None)
    state = lambda: fx.device_state().reshape(Q, gd[2], gd[1], gd[0])
    trace = refresh_trace(fx, state, 8 * md, md, chk)
    check_trace(trace, dim)
