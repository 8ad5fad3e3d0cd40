"""The headline shape on the device: D3Q19 256^3, FP64, Peskin4, md_per_lb 4,
the sphere of 1,810 Fibonacci markers plus 100,000 point particles (the
marker set of bench.py), driven through the fix.

- Against the oracle fix (OpenMP on the host cores) for two LB intervals from
  rest: populations within 1e-12 relative L2, per-atom forces within 1e-11
  (only the spreading's summation order differs), with the markers fixed and
  with the particles moving at every MD step.
- Size-independent properties over ten intervals on the device alone: mass
  conserved to FP64 reduction precision, and (without gravity) the fluid
  momentum gained in each interval equal to the reactions spread by its
  sub-steps (Guo's identity with the exact spreading normalisation).
- The reduced storage formats on the same workload against the FP64 path.
"""
import numpy as np
import pytest

import couplb_amd
from bench import make_markers
from tests.scenarios import (DeviceArray, atoms_from, copy_atoms, interior, rel_l2, run_device,
                             run_oracle)

pytestmark = pytest.mark.gpu

N, MD = 256, 4


def headline_atoms():
    x, v, m, _ = make_markers(N, 100000)
    return atoms_from(x, v, m)


def test_headline_shape_matches_oracle():
    cmd = f"flow all couplb {N} {N} {N} 0.1 1.0 kernel peskin4 md_per_lb {MD} gravity 1e-6 0.0 0.0"
    base = headline_atoms()
    print(f"case 256^3 FP64, {len(base['mass'])} markers, 2 LB intervals: device vs oracle", flush=True)
    fn = lambda step: copy_atoms(base)
    dev, fd = run_device(cmd, 2 * MD, dt=1.0 / MD, atoms_fn=fn)
    a = interior(dev.device_state().reshape(19, N + 2, N + 2, N + 2), 3)
    dev.close()
    ref, fo = run_oracle(cmd, 2 * MD, dt=1.0 / MD, atoms_fn=fn)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    del a
    ferr = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(fd, fo))
    print(f"  rel-L2(f) = {err:.3e}, max per-atom force difference / max = {ferr:.3e}", flush=True)
    assert err <= 1e-12
    assert ferr <= 1e-11
    ref.close()


def test_headline_shape_moving_markers_match_oracle():
    # the particles advance by v dt at every MD step (wrapped into the
    # periodic box), so every sub-step has new positions: the device orders
    # the markers, rebuilds the coverage and spreads node-centrically tile by
    # tile instead of using the fixed-marker pair map
    cmd = f"flow all couplb {N} {N} {N} 0.1 1.0 kernel peskin4 md_per_lb {MD} gravity 1e-6 0.0 0.0"
    base = headline_atoms()
    dt = 1.0 / MD

    def moved(step):
        a = copy_atoms(base)
        x = a["x"] + base["v"] * (dt * step)
        x -= N * np.floor(x / N)
        x[x >= N] = 0.0
        a["x"] = np.ascontiguousarray(x)
        return a

    print(f"case 256^3 FP64, {len(base['mass'])} moving markers, 2 LB intervals: device vs oracle",
          flush=True)
    dev, fd = run_device(cmd, 2 * MD, dt=dt, atoms_fn=moved)
    a = interior(dev.device_state().reshape(19, N + 2, N + 2, N + 2), 3)
    dev.close()
    ref, fo = run_oracle(cmd, 2 * MD, dt=dt, atoms_fn=moved)
    b = interior(ref.f(), 3)
    err = rel_l2(a, b)
    del a
    ferr = max(np.max(np.abs(x - y)) / np.max(np.abs(y)) for x, y in zip(fd, fo))
    print(f"  rel-L2(f) = {err:.3e}, max per-atom force difference / max = {ferr:.3e}", flush=True)
    assert err <= 1e-12
    assert ferr <= 1e-11
    ref.close()


def test_headline_shape_conserves_mass_and_momentum():
    cmd = f"flow all couplb {N} {N} {N} 0.1 1.0 kernel peskin4 md_per_lb {MD} check_every {MD}"
    fx = couplb_amd.Fix(cmd, dt=1.0 / MD)
    fx.init()
    a = headline_atoms()
    print(f"case 256^3 FP64, {len(a['mass'])} markers, 10 LB intervals: mass and momentum", flush=True)
    fx.setup(a)
    reaction = [-a["f"].sum(0)]      # setup's sub-step, spread into the first LB step
    diags = []
    for step in range(1, 10 * MD + 1):
        fx.step = step
        a["f"][:] = 0.0
        fx.post_force(a)
        fx.end_of_step(step)
        reaction.append(-a["f"].sum(0))
        if step % MD == 0:
            d = fx.last_diag()
            diags.append((step, d["mass"], d["mom"].copy(), d["clamps"]))
    fx.close()
    mass0 = float(N ** 3)
    for step, mass, mom, clamps in diags:
        assert clamps == 0
        assert abs(mass - mass0) <= 1e-12 * mass0, (step, mass)
    # the LB step at MD step s0 + 1 applies the reactions of MD sub-steps
    # s0 - MD + 1 .. s0 (entry 0: setup's), each scaled by 1 / md_per_lb
    # (force_scale = 1); check s1 = s0 + MD sees exactly that step
    for (s0, _, m0, _), (s1, _, m1, _) in zip(diags, diags[1:]):
        gained = m1 - m0
        spread = sum(reaction[s0 - MD + 1:s0 + 1]) / MD
        print(f"  steps {s0}->{s1}: fluid momentum gained {gained}, reactions {spread}", flush=True)
        assert np.allclose(gained, spread, rtol=1e-9, atol=1e-12 * np.abs(spread).max())


@pytest.mark.parametrize("storage", ["fp16", "bf16", "fp32"])
def test_headline_shape_reduced_storage_error(storage):
    # bench.py's workload (uniform flow 0.02, gravity, the 101,810 fixed
    # markers) for 20 LB steps with reduced storage against the FP64 device
    # path (tied to the oracle above): relative L2 of u within 1e-3 for FP16
    # (the target of the reduced formats), 2e-3 for BF16 (8 significant
    # bits), 1e-5 for FP32
    x, v, m, _ = make_markers(N, 100000)
    n = len(m)
    dx, dv, dm = DeviceArray(x), DeviceArray(v), DeviceArray(m)
    kw = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0, inv_N=1.0 / MD,
              force_scale=1.0)
    g = (1e-6, 0.0, 0.0)
    print(f"case 256^3 {storage} vs FP64, 20 LB steps", flush=True)
    fields = {}
    for st in (storage, "fp64"):
        d = couplb_amd.Device(3, N, N, N, tau=0.8, kernel=couplb_amd.PESKIN4, storage=st, md_per_lb=MD)
        d.init_equilibrium(1.0, (0.02, 0.0, 0.0))
        df = DeviceArray(np.zeros((n, 3)))
        for s in range(1 + 20 * MD):
            if s and (s - 1) % MD == 0:
                d.step(g, couplb_amd.RHO_REFRESHED_NEXT)
            d.ibm_device(n, dx.addr, dv.addr, dm.addr, df.addr, positions_unchanged=s > 0, **kw)
        d.sync()
        _, ux, uy, uz = d.macro()
        fields[st] = np.stack([ux, uy, uz])
        d.close()
        del df
    a, b = fields[storage], fields["fp64"]
    err = np.linalg.norm(a - b) / np.linalg.norm(b)
    print(f"  rel-L2(u) {err:.3e}", flush=True)
    bound = {"fp16": 1e-3, "bf16": 2e-3, "fp32": 1e-5}[storage]
    assert err <= bound
