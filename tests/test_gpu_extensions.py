"""Per-face wall velocities (keyword `wall_vel_face`), which the reference
lacks (one `wall_vel` for every moving face, fix_couplb.cpp:164-169, 532):
opposite moving walls on the device against the Couette closed form
u = U_lo + (U_hi - U_lo)(y + 1/2)/H and against the oracle."""
import numpy as np
import pytest

import oracle
from tests.scenarios import interior, run_device, run_oracle

pytestmark = pytest.mark.gpu


def _velocity(f, D):
    e, _ = oracle.lattice(D)
    e = np.asarray(e, float)             # (Q, 3)
    rho = f.sum(0)
    u = np.tensordot(e.T, f, axes=(1, 0)) / rho
    return rho, u


@pytest.mark.parametrize("D", [2, 3])
def test_opposite_moving_walls_couette(D):
    nz, dz = (4, 6) if D == 3 else (1, 1)
    cmd = (f"flow all couplb 4 16 {nz} 0.1 1.0 wall_y 2 2 "
           "wall_vel_face y lo -0.04 0 0 wall_vel_face y hi 0.04 0 0")
    steps = 20000
    print("case opposite walls D", D, flush=True)
    dev, _ = run_device(cmd, steps, dimension=D)
    ref, _ = run_oracle(cmd, steps, dimension=D)
    Q = 19 if D == 3 else 9
    shape = (Q, dz, 18, 6) if D == 3 else (Q, 1, 18, 6)
    fd = interior(dev.device_state().reshape(shape), D)
    fo = interior(ref.f(), D)
    assert np.array_equal(fd, fo)         # FP64 without markers: identical bits
    _, u = _velocity(fd, D)
    y = np.arange(16.0)
    exact = -0.04 + 0.08 * (y + 0.5) / 16
    err = np.max(np.abs(u[0] - exact[None, :, None]))
    print(f"D{D} Couette max |u - exact| = {err:.3e}")
    assert err < 1e-10
    assert np.max(np.abs(u[1:])) < 1e-12
