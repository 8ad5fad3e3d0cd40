"""Write the golden fixtures in tests/golden/*.npz.

Every expected output here is produced by the reference's own numerics
headers (CoupLB/couplb_lattice.h, couplb_collision.h, couplb_streaming.h,
couplb_boundary.h, couplb_ibm.h) compiled unmodified into
oracle/_ref/libcouplb_ref.so by oracle/Makefile, driven by the oracle's
restatement of FixCoupLB's call sequence (fix_couplb.cpp:551-901).
Inputs are stored next to the outputs so the tests replay them exactly.

  make -C oracle all ref && python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from tests.scenarios import atoms_from, copy_atoms, run_oracle  # noqa: E402


def ibm_atoms(n, L, seed):
    rng = np.random.default_rng(seed)
    return atoms_from(rng.uniform(0.0, L, (n, 3)), rng.normal(0.0, 1e-3, (n, 3)),
                      rng.uniform(0.5, 1.5, n))


def moving(base):
    def fn(step):
        a = copy_atoms(base)
        a["x"] = a["x"] + 0.05 * np.sin(0.3 * step + np.arange(len(a["mass"])))[:, None]
        return a
    return fn


CASES = {
    # C1 of the benchmark list: D2Q9 8x32 Poiseuille, 50000 steps
    "c1_poiseuille_2d": dict(
        cmd="flow all couplb 8 32 1 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 check_every 2000",
        steps=50000, dim=2, dt=1.0, atoms=None),
    # D2Q9 Couette: moving top wall (type 2), no-slip bottom
    "couette_2d": dict(
        cmd="flow all couplb 6 16 1 0.1 1.0 wall_y 1 2 wall_vel 0.02 0.0 0.0 check_every 500",
        steps=3000, dim=2, dt=1.0, atoms=None),
    # D3Q19, every wall type, Peskin4 IBM with moving markers, sub-stepping
    "walls_ibm_3d": dict(
        cmd="flow all couplb 8 6 7 0.1 1.0 wall_x 3 1 wall_y 2 4 wall_vel 0.01 0.0 0.005 "
            "gravity 1e-5 2e-6 -1e-6 kernel peskin4 md_per_lb 2 check_every 3",
        steps=20, dim=3, dt=0.5, atoms=(24, 6.0, 11)),
    # D3Q19 fully periodic, Roma IBM: stencils cross every periodic edge
    "periodic_ibm_3d": dict(
        cmd="flow all couplb 7 8 6 0.05 1.0 kernel roma md_per_lb 3 check_every 4",
        steps=15, dim=3, dt=1.0 / 3.0, atoms=(30, 6.0, 5)),
    # D2Q9 free-slip + open faces with IBM
    "slip_open_ibm_2d": dict(
        cmd="flow all couplb 10 9 1 0.08 1.0 wall_x 4 4 wall_y 3 1 gravity 2e-5 0.0 0.0 "
            "kernel peskin4 check_every 5",
        steps=40, dim=2, dt=1.0, atoms=(12, 9.0, 3)),
}


def run_case(name, c, numerics):
    atoms_fn = None
    if c["atoms"] is not None:
        n, L, seed = c["atoms"]
        base = ibm_atoms(n, L, seed)
        if c["dim"] == 2:
            base["x"][:, 2] = 0.0
            base["v"][:, 2] = 0.0
        atoms_fn = moving(base)
    fx, forces = run_oracle(c["cmd"], c["steps"], dimension=c["dim"], dt=c["dt"],
                            atoms_fn=atoms_fn, numerics=numerics)
    out = dict(f=fx.f(), rho=fx.field(oracle.F_RHO), ux=fx.field(oracle.F_UX),
               uy=fx.field(oracle.F_UY), uz=fx.field(oracle.F_UZ))
    d = fx.last_diag()
    out["diag"] = np.array([d["Ma"], d["mass"], *d["mom"], d["clamps"]])
    if forces:
        out["atom_forces"] = np.stack(forces)
    return out


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref not built: make -C oracle ref")
    for name, c in CASES.items():
        out = run_case(name, c, "ref")
        meta = np.array([c["cmd"], str(c["steps"]), str(c["dim"]), repr(c["dt"]), repr(c["atoms"])])
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), meta=meta, **out)
        print("wrote", name, {k: v.shape for k, v in out.items()})
    # delta kernels straight from the reference's delta_roma / delta_peskin4
    ops = oracle.ops("ref")
    r = np.linspace(-2.6, 2.6, 521)
    np.savez_compressed(os.path.join(HERE, "delta_kernels.npz"), r=r,
                        roma=np.array([ops.contents.delta(oracle.ROMA, x) for x in r]),
                        peskin4=np.array([ops.contents.delta(oracle.PESKIN4, x) for x in r]))
    print("wrote delta_kernels")


if __name__ == "__main__":
    main()
