"""C5 at its full shape: D3Q19 256^3 shear flow between walls moving at
-/+0.04, 64 fixed spheres of diameter 16 (804 Peskin4 markers each, 51,456 in
all), multi-direct forcing (CG) to slip 1e-6, the warm-start network trained
on the GPU with the solver's own converged forces.

Every solve must converge; solves started from zero, from the network and
from the previous step's forces must give the same per-sphere forces (to the
1e-3 a slip tolerance leaves them, see tests/test_mdf_locality.py) and the
same next fluid state; the previous-step start must cut the passes, and the
network start must not cost more than the start from zero.  Block-Jacobi
preconditioned CG must reach the same forces and state in fewer passes than
plain CG: at most 0.85x from zero and 0.7x from the previous forces.  FP16 storage
with the network start stays within 1e-3 (relative L2 of u) of the FP64
solve from zero."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 256
WALLS = ((0, 0), (2, 2), (0, 0))
FACE = [[[0, 0, 0], [0, 0, 0]], [[-0.04, 0, 0], [0.04, 0, 0]], [[0, 0, 0], [0, 0, 0]]]
KW = dict(lo=(0, 0, 0), inv_dx=1.0, vel_scale=1.0, force_scale=1.0, tol=1e-6, max_iter=4000,
          method="cg")
TRAIN, HELD = 20, 6


def fib(n, R, c):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1) * R + c


def spheres(count=64, D=16.0, seed=2026):
    """seeded non-overlapping centres, at least D/2 + 2 from the y walls"""
    rng = np.random.default_rng(seed)
    cs = []
    while len(cs) < count:
        c = np.array([rng.uniform(0, N), rng.uniform(D / 2 + 2, N - D / 2 - 2), rng.uniform(0, N)])
        if all(np.linalg.norm(np.minimum(np.abs(c - o), [N, 1e9, N] - np.abs(c - o))) >= D + 2 for o in cs):
            cs.append(c)
    ns = int(round(math.pi * D * D))
    return np.concatenate([fib(ns, D / 2, c) % [N, 1e9, N] for c in cs]), ns


def couette():
    import oracle
    y = np.arange(-1, N + 1) + 0.5
    prof = -0.04 + 0.08 * y / N
    fy = oracle.feq(3, np.ones((N + 2, 1)), np.stack([prof[:, None], np.zeros((N + 2, 1)),
                                                      np.zeros((N + 2, 1))]))
    return np.ascontiguousarray(np.broadcast_to(fy[:, None, :, :], (19, N + 2, N + 2, N + 2)))


def test_c5_mdf_with_the_trained_network():
    # device buffers through the library's own HIP runtime (DeviceArray)
    import couplb_amd
    from tests.scenarios import DeviceArray
    x, ns = spheres()
    n = len(x)
    tx = DeviceArray(x)
    tV = DeviceArray(np.zeros((n, 3)))

    def device(storage="fp64"):
        return couplb_amd.Device(3, N, N, N, tau=0.8, walls=WALLS, face_vel=FACE,
                                 kernel=couplb_amd.PESKIN4, storage=storage)

    print("case C5 256^3, 64 spheres: training data", flush=True)
    A = device()
    A.set_state(couette())
    feats = DeviceArray(np.zeros((TRAIN * n, 67)), np.float32)
    targs = DeviceArray(np.zeros((TRAIN * n, 3)), np.float32)
    fb = DeviceArray(np.zeros((n, 3)))
    for s in range(TRAIN):
        st = A.mdf_device(n, tx.addr, tV.addr, None, fb.addr, positions_unchanged=s > 0,
                          features_ptr=feats.addr + s * n * 67 * 4, targets_ptr=targs.addr + s * n * 3 * 4,
                          **KW)
        assert st["converged"]
        A.step((0.0, 0.0, 0.0), 0)
    A.sync()
    net = couplb_amd.Mlp(seed=11)
    l0, _ = net.gradient_device(TRAIN * n, feats.addr, targs.addr)
    net.train_device(TRAIN * n, feats.addr, targs.addr, batch=4096, epochs=8, lr=2e-3, seed=5)
    l1, _ = net.gradient_device(TRAIN * n, feats.addr, targs.addr)
    print(f"  network loss {l0:.4g} -> {l1:.4g} on {TRAIN * n} samples", flush=True)
    assert l1 < l0
    start = A.get_state()
    A.close()

    def run(mode, storage="fp64", method="cg"):
        print("case C5 held-out steps:", mode, storage, method, flush=True)
        d = device(storage)
        d.set_state(start)
        f = DeviceArray(np.zeros((n, 3)))
        fprev = DeviceArray(np.zeros((n, 3)))
        passes, totals = [], []
        for s in range(HELD):
            st = d.mdf_device(n, tx.addr, tV.addr, fprev.addr if (mode == "previous" and s > 0) else None,
                              f.addr, positions_unchanged=s > 0,
                              warm_start=net if mode == "network" else None, **dict(KW, method=method))
            d.sync()
            assert st["converged"] and st["slip_last"] < 1e-6
            passes.append(st["iterations"])
            fh = f.download()
            totals.append(fh.reshape(64, ns, 3).sum(1))
            fprev.upload(-fh)          # F (lattice units) = -f_body with force_scale 1
            d.step((0.0, 0.0, 0.0), 0)
        d.sync()
        u = np.stack(d.macro()[1:])
        d.close()
        return np.array(passes), np.array(totals), u

    p0, t0, u0 = run("zero")
    pn, tn, un = run("network")
    pp, tp, up = run("previous")
    print(f"  passes: zero {p0.tolist()} network {pn.tolist()} previous {pp.tolist()}")
    for t, u in ((tn, un), (tp, up)):
        assert np.max(np.abs(t - t0)) <= 1e-2 * np.max(np.abs(t0))
        assert np.linalg.norm(u - u0) <= 1e-4 * np.linalg.norm(u0)
    assert pp[1:].mean() <= 0.8 * p0[1:].mean()
    assert pn.mean() <= 1.05 * p0.mean()
    # block-Jacobi preconditioned CG: the same forces and state in fewer passes
    pz, tz, uz = run("zero", method="pcg")
    pq, tq, uq = run("previous", method="pcg")
    print(f"  pcg passes: zero {pz.tolist()} previous {pq.tolist()}")
    for t, u in ((tz, uz), (tq, uq)):
        assert np.max(np.abs(t - t0)) <= 1e-2 * np.max(np.abs(t0))
        assert np.linalg.norm(u - u0) <= 1e-4 * np.linalg.norm(u0)
    # measured: 0.77x the passes of plain CG from zero, 0.58x from the
    # previous forces
    assert pz.mean() <= 0.85 * p0.mean()
    assert pq[1:].mean() <= 0.7 * pp[1:].mean()
    _, _, u16 = run("network", "fp16")
    err = np.linalg.norm(u16 - u0) / np.linalg.norm(u0)
    print(f"  fp16 + network vs fp64 from zero: rel-L2(u) {err:.3e}")
    assert err <= 1e-3
