#!/usr/bin/env python3
"""D3Q19 LB step with immersed-boundary coupling on one MI355X.

Workload (synthetic, seeded): D3Q19 256^3, fully periodic, nu = 0.1
(tau = 0.8), FP64 populations, uniform initial velocity (0.02, 0, 0) with a
constant body force g_x = 1e-6; one rigid sphere of diameter 24 dx at the box
centre (1810 Fibonacci-lattice markers, held fixed, v = 0) plus 100000 point
particles at uniform random positions with Gaussian velocities
(sigma = 1e-3), all of mass 1; Peskin4 kernel, xi_ibm = 1, md_per_lb = 4.
Positions are prescribed and held fixed, velocities constant (`value`);
`mlups_markers_moving` is the same workload with the particles advanced by
v * dt_MD at every MD sub-step.

A step = one LB step (the fused pull kernel) + md_per_lb IBM sub-steps, with
the marker arrays resident in HBM.  value = MLUPS = 256^3 * steps / wall
seconds / 1e6.

  python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import gc
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# algorithmic bytes of one D3Q19 Peskin4 FP64 marker sub-step: interpolate
# 64 nodes x 3 x 8 B, spread 64 x 3 x 8 B, marker I/O x 24 + v 24 + m 8 + f RMW 48
IBM_BYTES_PER_MARKER = 64 * 3 * 8 * 2 + 104
Q = 19


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def fibonacci_sphere(n, radius, centre):
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1) * radius + np.asarray(centre)


def make_markers(N, n_particles, seed=2025):
    D = 24.0
    ns = int(round(math.pi * D * D))  # ~1 dx marker spacing: 1810
    sphere = fibonacci_sphere(ns, D / 2, (N / 2, N / 2, N / 2))
    rng = np.random.default_rng(seed)
    px = rng.uniform(0.0, float(N), (n_particles, 3))
    pv = rng.normal(0.0, 1e-3, (n_particles, 3))
    x = np.concatenate([sphere, px]).astype(np.float64)
    v = np.concatenate([np.zeros_like(sphere), pv]).astype(np.float64)
    m = np.ones(len(x))
    return x, v, m, ns


def cpu_baseline(N, nz_sample, steps, md_per_lb):
    """Reference-numerics LB/IBM cycle on host cores (bounded sample)."""
    import oracle
    kind = "reference"
    try:
        oracle.ops("ref")
        numerics = "ref"
    except OSError:
        numerics, kind = "oracle", "port"
    x, v, m, _ = make_markers(N, 100000)
    keep = x[:, 2] < nz_sample                  # markers inside the slab
    x, v, m = x[keep], v[keep], m[keep]
    fx = oracle.Fix(N, N, nz_sample, 0.1, 1.0, md_per_lb=md_per_lb, dt=1.0 / md_per_lb,
                    gravity=(1e-6, 0.0, 0.0), kernel=oracle.PESKIN4, numerics=numerics)
    n = len(m)
    atoms = dict(x=x, v=v, f=np.zeros((n, 3)), mask=np.ones(n, np.int32), mass=m)
    fx.setup(atoms)
    t0 = time.perf_counter()
    for _ in range(steps * md_per_lb):
        atoms["f"][:] = 0
        fx.post_force(atoms)
    dt = time.perf_counter() - t0
    threads = oracle.omp_threads()
    nodes = N * N * nz_sample
    return dict(value=nodes * steps / dt / 1e6, unit="MLUPS", cores=threads, kind=kind,
                sample=(f"D3Q19 {N}x{N}x{nz_sample} slab of the same workload, {n} markers "
                        f"(same density), md_per_lb={md_per_lb}, {steps} LB steps, FP64, "
                        f"reference pass structure (numerics: {numerics}), "
                        f"OMP_NUM_THREADS={threads}, {dt:.1f} s"))


def host_cpu():
    """CPU model and logical CPU count of this host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def copy_calibration(mb, reps):
    """Plain device-to-device copy of known size: the reference point for
    rocprofv3 FETCH_SIZE / WRITE_SIZE and for achievable HBM bandwidth."""
    import torch
    n = mb * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    print(json.dumps({"copy_mb": mb, "bytes_per_copy": 2 * n * 8, "avg_ms": ms,
                      "copy_gbs": 2 * n * 8 / (ms * 1e-3) / 1e9}), flush=True)


def storage_error(N, storage, md, tx, tv, tm, ibm_kw, g, steps=20):
    """Velocity error of the workload run with `storage` against the FP64
    device path (tied to the oracle by the parity tests), both from the same
    state and markers after `steps` LB steps: rel-L2 and rel-Linf of u, and
    rel-L2 of the disturbance u - u0."""
    import torch
    import couplb_amd
    n = tm.numel()
    fields = {}
    for st in (storage, "fp64"):
        d = couplb_amd.Device(3, N, N, N, tau=0.8, kernel=couplb_amd.PESKIN4, storage=st, md_per_lb=md)
        d.init_equilibrium(1.0, (0.02, 0.0, 0.0))
        tf = torch.zeros((n, 3), dtype=torch.float64, device="cuda:0")
        torch.cuda.synchronize()
        for s in range(1 + steps * md):
            if s and (s - 1) % md == 0:
                d.step(g, couplb_amd.RHO_REFRESHED_NEXT)
            d.ibm_device(n, tx.data_ptr(), tv.data_ptr(), tm.data_ptr(), tf.data_ptr(),
                         positions_unchanged=s > 0, **ibm_kw)
        d.sync()
        _, ux, uy, uz = d.macro()
        fields[st] = np.stack([ux, uy, uz])
        d.close()
        del tf
    a, b = fields[storage], fields["fp64"]
    u0 = np.array([0.02, 0.0, 0.0]).reshape(3, 1, 1, 1)
    return dict(steps=steps, rel_l2_u=float(np.linalg.norm(a - b) / np.linalg.norm(b)),
                rel_linf_u=float(np.max(np.abs(a - b)) / np.max(np.abs(b))),
                rel_l2_disturbance=float(np.linalg.norm(a - b) / np.linalg.norm(b - u0)))


def pmc_traffic(storage, n):
    """HBM bytes per k_step launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (scaled by a plain copy of known size), read from the
    committed summary of that measurement; None when there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"pmc_k_step_d3q19_{n}cube_{storage}.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None



# ---------------------------------------------------------------------------
# Further workloads (not the headline line): python bench.py --workload c1|c2|c4|c5
# ---------------------------------------------------------------------------
D3Q19_E = np.array([[0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
                    [0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1],
                    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1]], float).T
D3Q19_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def feq_field(rho, u):
    """D3Q19 equilibrium w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u)
    (couplb_lattice.h:137-144) for fields rho (...), u (3, ...), velocities
    in the reference's order (couplb_lattice.h:49-57): returns (19, ...)."""
    eu = np.tensordot(D3Q19_E, u, axes=(1, 0))
    uu = (u * u).sum(0)
    w = D3Q19_W.reshape((19,) + (1,) * np.ndim(uu))
    return w * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * uu)


def ghost_coords(n):
    """cell-centre coordinate of ghost-inclusive index 0..n+1 (x = i - 1 + 1/2)"""
    return np.arange(-1, n + 1) + 0.5


def workload_c1():
    """D2Q9 8x32 Poiseuille channel (the reference's own CPU-runnable deck:
    walls in y, g_x = 5e-6, nu = 0.1, rest start, 50,000 steps).  A step of
    256 nodes is far shorter than a kernel launch: the run of steps goes
    through couplb_lbm_steps_async, which keeps the state in LDS and runs the
    steps inside one launch; one launch per step and the fix driven step by
    step from Python are timed beside it, and the reference's pass
    structure on host cores as the baseline."""
    import couplb_amd
    import oracle
    steps, g = 50000, (5e-6, 0.0, 0.0)
    walls = ((0, 0), (1, 1), (0, 0))
    exact = 5e-6 / (2 * 0.1) * (np.arange(32) + 0.5) * (32 - 0.5 - np.arange(32))
    out = {}
    for name, bound in (("fused", None), ("launch_per_step", "0")):
        if bound is not None:
            os.environ["COUPLB_FUSED_NODES"] = bound
        d = couplb_amd.Device(2, 8, 32, 1, tau=0.8, walls=walls)
        os.environ.pop("COUPLB_FUSED_NODES", None)
        d.init_equilibrium(1.0)
        d.sync()
        t0 = time.perf_counter()
        d.steps_async(steps, g, 0)
        d.sync()
        dt = time.perf_counter() - t0
        rho, ux, uy, uz = d.macro()
        prof = ux.reshape(32, 8).mean(axis=1)
        out[name] = dict(mlups=round(256 * steps / dt / 1e6, 1), us_per_step=round(dt / steps * 1e6, 3),
                         fuses=d.fuses_steps(),
                         profile_max_rel_err=float(np.max(np.abs(prof - exact)) / exact.max()))
        d.close()
    fx = couplb_amd.Fix("flow all couplb 8 32 1 0.1 1.0 wall_y 1 1 gravity 5e-6 0.0 0.0 check_every 2000",
                        dimension=2, dt=1.0)
    fx.init()
    fx.setup(None)
    t0 = time.perf_counter()
    fx.run(steps)
    couplb_amd.lib().couplb_synchronize(couplb_amd.lib().couplb_fix_context(fx.h))
    dt = time.perf_counter() - t0
    out["fix_python_driver"] = dict(mlups=round(256 * steps / dt / 1e6, 1),
                                    us_per_step=round(dt / steps * 1e6, 3))
    fx.close()
    cpu = None
    try:
        fo = oracle.Fix(8, 32, 1, 0.1, 1.0, dim=2, walls=walls, gravity=g, numerics="ref")
        kind = "reference"
    except OSError:
        fo = oracle.Fix(8, 32, 1, 0.1, 1.0, dim=2, walls=walls, gravity=g)
        kind = "port"
    fo.setup(None)
    dt = fo.time_lbm_steps(steps, g)
    fo.close()
    cpu = dict(value=round(256 * steps / dt / 1e6, 2), unit="MLUPS", cores=oracle.omp_threads(), kind=kind,
               sample=f"the whole workload: 50,000 steps, reference pass structure, {dt:.2f} s")
    return dict(metric="D2Q9 8x32 channel MLUPS (steps after the first in one launch)", unit="MLUPS",
                value=out["fused"]["mlups"], steps=steps, ms_per_step=round(1e3 / (out["fused"]["mlups"] * 1e6 / 256), 6),
                data="synthetic (rest start)",
                config={"workload": "D2Q9 8x32 channel, no-slip walls in y, g_x=5e-6, nu=0.1, rest start, "
                                    "50,000 steps"},
                results=out, cpu_baseline=cpu)


def workload_c2(steps):
    """D3Q19 channel 128^3, walls in y, g_x = 1e-6, initial Poiseuille + seeded
    uniform noise 1e-4: throughput and profile at FP64 and FP32."""
    import couplb_amd
    n, nu, g = 128, 0.1, 1e-6
    y = ghost_coords(n)
    prof = g / (2 * nu) * (y + 0.0) * (n - y)           # u(y) with walls at y = 0, n
    rng = np.random.default_rng(1234)
    shape = (n + 2, n + 2, n + 2)
    u = np.zeros((3,) + shape)
    u[0] = prof[None, :, None]
    u += rng.uniform(-1e-4, 1e-4, u.shape)
    f0 = feq_field(np.ones(shape), u)
    del u
    out = {}
    fields = {}
    exact = g / (2 * nu) * (np.arange(n) + 0.5) * (n - 0.5 - np.arange(n))
    for storage in ("fp64", "fp32"):
        d = couplb_amd.Device(3, n, n, n, tau=3 * nu + 0.5, walls=((0, 0), (1, 1), (0, 0)),
                              storage=storage)
        d.set_state(f0)
        d.sync()
        t0 = time.perf_counter()
        d.steps_async(steps, (g, 0.0, 0.0), 0)
        d.sync()
        dt = time.perf_counter() - t0
        rho, ux, uy, uz = d.macro()
        fields[storage] = np.stack([ux, uy, uz])
        px = ux.mean(axis=(0, 2))
        out[storage] = dict(mlups=round(n ** 3 * steps / dt / 1e6, 1),
                            profile_max_rel_err=float(np.max(np.abs(px - exact)) / exact.max()),
                            mass=float(rho.sum()))
        d.close()
    a, b = fields["fp32"], fields["fp64"]
    out["fp32"]["rel_l2_u_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["fp32"]["rel_linf_u_vs_fp64"] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return dict(metric="D3Q19 128^3 channel MLUPS and profile error", unit="MLUPS",
                value=out["fp64"]["mlups"], steps=steps, data="synthetic (seeded noise)",
                config={"workload": "D3Q19 128^3 channel, no-slip walls in y, g_x=1e-6, nu=0.1, "
                                    "initial Poiseuille + uniform noise 1e-4"}, results=out)


def workload_c4(steps, n=384):
    """Taylor-Green vortex n^3 (z-invariant), u0 = 0.03, k = 2 pi / n,
    nu = 0.02: FP64 against FP32 / FP16 / BF16 deviation storage and the
    16-bit fixed point (range 0.0075: the largest |f - w rho0| of this
    vortex is 0.00515)."""
    import couplb_amd
    nu, u0 = 0.02, 0.03
    k = 2 * np.pi / n
    c = ghost_coords(n)
    X, Y = np.meshgrid(c, c, indexing="xy")          # (gy, gx)
    ux = u0 * np.sin(k * X) * np.cos(k * Y)
    uy = -u0 * np.cos(k * X) * np.sin(k * Y)
    rho = 1.0 + 0.75 * u0 * u0 * (np.cos(2 * k * X) + np.cos(2 * k * Y))
    f2 = feq_field(rho, np.stack([ux, uy, np.zeros_like(ux)]))          # (19, gy, gx)
    f0 = np.ascontiguousarray(np.broadcast_to(f2[:, None], (19, n + 2, n + 2, n + 2)))
    del f2
    decay = float(np.exp(-4 * nu * k * k * steps))   # KE(t)/KE(0), analytic
    res, ref = {}, None

    def ke(d):
        r, a, b, cz = d.macro()
        e = 0.5 * float((r * (a * a + b * b + cz * cz)).sum())
        return e, np.stack([a, b, cz])

    for storage in ("fp64", "fp32", "fp16", "bf16", "i16"):
        d = couplb_amd.Device(3, n, n, n, tau=3 * nu + 0.5, storage=storage, storage_range=0.0075)
        d.set_state(f0)
        m0 = d.diagnostics()["mass"]
        e0, _ = ke(d)
        d.sync()
        t0 = time.perf_counter()
        d.steps_async(steps, (0.0, 0.0, 0.0), 0)
        d.sync()
        dt = time.perf_counter() - t0
        e1, u = ke(d)
        mass = d.diagnostics()["mass"]
        d.close()
        r = dict(mlups=round(n ** 3 * steps / dt / 1e6, 1), ke_ratio=e1 / e0,
                 ke_ratio_rel_err_vs_analytic=(e1 / e0) / decay - 1.0, mass=mass, mass_drift=mass / m0 - 1.0)
        if ref is None:
            ref = u
        else:
            r["rel_l2_u_vs_fp64"] = float(np.linalg.norm(u - ref) / np.linalg.norm(ref))
        res[storage] = r
        del u
    return dict(metric=f"TGV {n}^3 precision study: KE decay and rel-L2(u) vs FP64 GPU",
                unit="MLUPS", value=res["fp16"]["mlups"], steps=steps,
                data="synthetic (analytic initial field)",
                config={"workload": f"D3Q19 {n}^3 periodic Taylor-Green, u0=0.03, k=2pi/{n}, nu=0.02",
                        "analytic_ke_ratio": decay}, results=res)


def c5_spheres(N, count=64, D=16.0, seed=2026):
    """Non-overlapping sphere centres at least D/2 + 2 from the y walls."""
    rng = np.random.default_rng(seed)
    cs = []
    while len(cs) < count:
        c = np.array([rng.uniform(0, N), rng.uniform(D / 2 + 2, N - D / 2 - 2), rng.uniform(0, N)])
        ok = True
        for o in cs:
            d = np.abs(c - o)
            d[0] = min(d[0], N - d[0])
            d[2] = min(d[2], N - d[2])
            if np.linalg.norm(d) < D + 2:
                ok = False
                break
        if ok:
            cs.append(c)
    ns = int(round(math.pi * D * D))                      # 804 markers, ~1 dx spacing
    x = np.concatenate([fibonacci_sphere(ns, D / 2, c) % [N, 1e9, N] for c in cs])
    return x, ns


def workload_c5(steps, train_steps=20, N=256):
    """Shear flow between walls moving at -/+0.04 with 64 fixed spheres of
    diameter 16 (Peskin4 markers), multi-direct forcing to slip 1e-6 (CG,
    plain and block-Jacobi preconditioned).  Compares the solve started from
    zero (FP64, plain), from the previous step's forces, and from the trained
    network (FP64 and FP16 storage); the value is FP16 storage with the
    network start and the preconditioned solve."""
    import torch
    import couplb_amd
    cuda = torch.device("cuda:0")
    walls = ((0, 0), (2, 2), (0, 0))
    face = [[[0, 0, 0], [0, 0, 0]], [[-0.04, 0, 0], [0.04, 0, 0]], [[0, 0, 0], [0, 0, 0]]]
    x, ns = c5_spheres(N)
    n = len(x)
    tx = torch.from_numpy(x).to(cuda)
    tV = torch.zeros((n, 3), dtype=torch.float64, device=cuda)
    tf = torch.zeros((n, 3), dtype=torch.float64, device=cuda)
    kw = dict(lo=(0, 0, 0), inv_dx=1.0, vel_scale=1.0, force_scale=1.0, tol=1e-6, max_iter=4000,
              method="cg")
    # initial Couette profile between the walls
    y = ghost_coords(N)
    prof = -0.04 + 0.08 * y / N
    fy = feq_field(np.ones((N + 2, 1)), np.stack([prof[:, None], np.zeros((N + 2, 1)),
                                                  np.zeros((N + 2, 1))]))      # (19, gy, 1)
    f0 = np.ascontiguousarray(np.broadcast_to(fy[:, None, :, :], (19, N + 2, N + 2, N + 2)))
    del fy

    def device(storage):
        d = couplb_amd.Device(3, N, N, N, tau=0.8, walls=walls, face_vel=face,
                              kernel=couplb_amd.PESKIN4, storage=storage)
        return d

    A = device("fp64")
    A.set_state(f0)
    del f0
    # training data: the solver's converged forces over the first train_steps LB steps
    feats = torch.zeros((train_steps * n, 67), dtype=torch.float32, device=cuda)
    targs = torch.zeros((train_steps * n, 3), dtype=torch.float32, device=cuda)
    t0 = time.perf_counter()
    it_tr = []
    for s in range(train_steps):
        st = A.mdf_device(n, tx.data_ptr(), tV.data_ptr(), None, tf.data_ptr(),
                          positions_unchanged=s > 0, features_ptr=feats[s * n].data_ptr(),
                          targets_ptr=targs[s * n].data_ptr(), **kw)
        it_tr.append(st["iterations"])
        A.step((0.0, 0.0, 0.0), 0)
    A.sync()
    gen_s = time.perf_counter() - t0
    net = couplb_amd.Mlp(seed=11)
    t0 = time.perf_counter()
    loss0, _ = net.gradient_device(len(feats), feats.data_ptr(), targs.data_ptr())
    lossE = net.train_device(len(feats), feats.data_ptr(), targs.data_ptr(), batch=4096, epochs=8,
                             lr=2e-3, seed=5)
    loss1, _ = net.gradient_device(len(feats), feats.data_ptr(), targs.data_ptr())
    train_s = time.perf_counter() - t0
    start = A.get_state()
    A.close()

    def run(storage, mode, method="cg"):
        d = device(storage)
        d.set_state(start)
        d.sync()
        iters, F0 = [], None
        fprev = torch.zeros_like(tf)
        fprev.copy_(-tf)            # load torch's kernels before the clock starts
        fprev.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for s in range(steps):
            st = d.mdf_device(n, tx.data_ptr(), tV.data_ptr(),
                              fprev.data_ptr() if (mode == "previous" and s > 0) else None,
                              tf.data_ptr(), positions_unchanged=s > 0,
                              warm_start=net if mode == "network" else None, **dict(kw, method=method))
            iters.append(st["iterations"])
            if mode == "previous":
                d.sync()                  # the device stream does not order with torch's
                fprev.copy_(-tf)          # F (lattice) = -f_body with force_scale 1
                torch.cuda.synchronize()
            d.step((0.0, 0.0, 0.0), 0)
        d.sync()
        wall = time.perf_counter() - t
        r, a, b, c = d.macro()
        d.close()
        return dict(mlups=round(N ** 3 * steps / wall / 1e6, 1), mean_iterations=float(np.mean(iters)),
                    iterations=iters), np.stack([a, b, c])

    res = {}
    res["fp64_plain"], u_ref = run("fp64", "zero")
    # "pcg": conjugate gradients preconditioned by 32-marker blocks
    for name, storage, mode, method in (("fp64_previous", "fp64", "previous", "cg"),
                                        ("fp64_network", "fp64", "network", "cg"),
                                        ("fp16_network", "fp16", "network", "cg"),
                                        ("fp16_plain", "fp16", "zero", "cg"),
                                        ("fp64_pcg", "fp64", "zero", "pcg"),
                                        ("fp64_pcg_previous", "fp64", "previous", "pcg"),
                                        ("fp16_network_pcg", "fp16", "network", "pcg"),
                                        ("fp16_pcg_previous", "fp16", "previous", "pcg")):
        r, u = run(storage, mode, method)
        r["rel_l2_u_vs_fp64_plain"] = float(np.linalg.norm(u - u_ref) / np.linalg.norm(u_ref))
        res[name] = r
        del u
    return dict(metric="C5 learned MDF warm start: MLUPS, MDF passes, rel-L2(u) vs FP64 plain",
                unit="MLUPS", value=res["fp16_network_pcg"]["mlups"], steps=steps,
                data="synthetic (seeded sphere centres, Couette initial field)",
                config={"workload": f"D3Q19 {N}^3 shear flow, walls y at -/+0.04, 64 fixed spheres "
                                    f"D=16 ({ns} markers each, {n} total), Peskin4, MDF (CG, plain and "
                                    f"block-Jacobi preconditioned) to "
                                    f"slip 1e-6, network 67-64-64-3 trained on {len(feats)} "
                                    f"samples from {train_steps} solves"},
                training=dict(samples=len(feats), generate_s=round(gen_s, 2), train_s=round(train_s, 2),
                              loss_initial=loss0, loss_last_epoch=lossE, loss_final=loss1,
                              train_phase_iterations=it_tr),
                results=res)

def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--n", type=int, default=256, help="grid edge (interior nodes)")
    ap.add_argument("--particles", type=int, default=100000)
    ap.add_argument("--storage", default="fp64", choices=["fp64", "fp32", "fp16", "bf16", "i16"])
    ap.add_argument("--cpu-steps", type=int, default=40)
    ap.add_argument("--cpu-nz", type=int, default=64)
    ap.add_argument("--cpu-only", default="", metavar="NZ,STEPS",
                    help="run only the CPU baseline on an NZ-plane slab for STEPS LB steps "
                         "and print it (no GPU; used for the one-thread figure)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                    help="time a plain device copy of MB megabytes (PMC byte-counter scale)")
    ap.add_argument("--workload", default="c3", choices=["c3", "c1", "c2", "c4", "c5"],
                    help="c3 (default): the headline line; c1 small 2D channel, c2 channel, "
                         "c4 precision study, c5 learned MDF warm start")
    ap.add_argument("--grid", type=int, default=0, help="grid edge for c4 (default 384)")
    args = ap.parse_args()
    # the timed regions run without Python's cyclic garbage collector (as
    # timeit does): a full collection inside a timed region blocked the
    # host for ~40 ms between two IBM calls.  Everything here is freed by
    # reference counting.
    gc.collect()
    gc.disable()
    if args.cpu_only:
        nz, steps = (int(t) for t in args.cpu_only.split(","))
        print(json.dumps(cpu_baseline(args.n, nz, steps, 4)), flush=True)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: one GPU is measured; multi-GPU runs are not part of it "
              f"(asked for --gpus {args.gpus}, WORLD_SIZE={world})", file=sys.stderr)
        sys.exit(2)

    import torch
    import couplb_amd

    if args.copy_calibration:
        copy_calibration(args.copy_calibration, args.steps)
        return
    if args.workload != "c3":
        if args.workload == "c1":
            out = workload_c1()
        elif args.workload == "c2":
            out = workload_c2(args.steps)
        elif args.workload == "c4":
            out = workload_c4(args.steps, args.grid or 384)
        else:
            out = workload_c5(args.steps)
        out.update(n_gpus=1, higher_is_better=True, dtype="f64 / reduced storage as named")
        print(json.dumps(out), flush=True)
        return
    N, md = args.n, 4
    log(f"bench: D3Q19 {N}^3 {args.storage} IBM md_per_lb={md} particles={args.particles}")
    dev = couplb_amd.Device(3, N, N, N, tau=0.8, kernel=couplb_amd.PESKIN4,
                            storage=args.storage, md_per_lb=md)
    dev.init_equilibrium(1.0, (0.02, 0.0, 0.0))
    x, v, m, nsphere = make_markers(N, args.particles)
    n = len(m)
    cuda = torch.device("cuda:0")
    tx = torch.from_numpy(x).to(cuda)
    tv = torch.from_numpy(v).to(cuda)
    tm = torch.from_numpy(m).to(cuda)
    tf = torch.zeros((n, 3), dtype=torch.float64, device=cuda)
    torch.cuda.synchronize()
    ibm_kw = dict(lo=(0.0, 0.0, 0.0), inv_dx=1.0, xi_over_dtLB=1.0, vel_scale=1.0,
                  inv_N=1.0 / md, force_scale=1.0)
    g = (1e-6, 0.0, 0.0)

    # The workload's markers are prescribed and fixed, so every sub-step but
    # the very first passes positions_unchanged.  moving=True re-declares the
    # positions as new at every sub-step (the path of moving particles).
    def ibm(first=False, moving=False):
        dev.ibm_device(n, tx.data_ptr(), tv.data_ptr(), tm.data_ptr(), tf.data_ptr(),
                       positions_unchanged=not (first or moving), **ibm_kw)

    ibm(first=True)  # FixCoupLB::setup primes the IBM force (fix_couplb.cpp:558)

    # every call is asynchronous on the context stream; kernel times come
    # from HIP events on that stream, read back after the timed region
    def one_step(moving=False):
        dev.step(g, couplb_amd.RHO_REFRESHED_NEXT)
        for _ in range(md):
            ibm(moving=moving)

    for _ in range(args.warmup):
        one_step()
    dev.sync()
    torch.cuda.synchronize()
    dev.kernel_ms_reset()
    # inside the timed region only the LB kernel (the roofline's) records
    # events; the IBM breakdown comes from a second pass below
    dev.set_timer_mask(1 << 0)
    t0 = time.perf_counter()
    for s in range(args.steps):
        one_step()
        if (s + 1) % 10 == 0:
            log(f"bench: step {s + 1}/{args.steps} issued")
    dev.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lb_tot, lb_n = dev.kernel_ms_total(0)
    # the same steps again with every region timed: IBM sub-step, spreading,
    # velocity gather (the events between those kernels cost about 3 % of a
    # step, which is why they stay out of the timed region)
    dev.set_timer_mask(0xFFFFFFFF)
    dev.kernel_ms_reset()
    for _ in range(args.steps):
        one_step()
    dev.sync()
    ibm_tot, ibm_n = dev.kernel_ms_total(1)
    sp_tot, sp_n = dev.kernel_ms_total(2)
    ga_tot, ga_n = dev.kernel_ms_total(4)
    nodes = N ** 3
    mlups = nodes * args.steps / wall / 1e6
    s_bytes = {"fp64": 8, "fp32": 4, "fp16": 2, "bf16": 2, "i16": 2}[args.storage]
    b_lup = 2 * Q * s_bytes
    lb_avg = lb_tot / max(lb_n, 1)
    ibm_avg = ibm_tot / max(ibm_n, 1)
    spread_avg = sp_tot / max(args.steps, 1)
    achieved = b_lup * nodes / (lb_avg * 1e-3) / 1e9 if lb_avg > 0 else 0.0
    diag = dev.diagnostics()

    # LB-only throughput of the same grid (no markers), for reference
    dev2_steps = max(10, args.steps)
    dev.sync()
    t1 = time.perf_counter()
    dev.steps_async(dev2_steps, g, 0)
    dev.sync()
    lb_only = nodes * dev2_steps / (time.perf_counter() - t1) / 1e6
    # the LB kernel alone (no spread force), timed by events: the fixed-marker
    # map's force sums run inside the LB kernel, and the difference is their
    # share of it
    dev.set_timer_mask(1 << 0)
    dev.kernel_ms_reset()
    for _ in range(10):
        dev.step(g, 0)
    dev.sync()
    plain_tot, plain_n = dev.kernel_ms_total(0)
    lb_plain = plain_tot / max(plain_n, 1)
    spread_in_lb = max(0.0, lb_avg - lb_plain)

    # IBM time per marker sub-step: the sub-step kernels plus this sub-step's
    # share of the force spreading that precedes the next LB kernel
    spread_lb = spread_avg + spread_in_lb
    markers_per_s = n / ((ibm_avg + spread_lb / md) * 1e-3) if n and ibm_avg > 0 else 0.0
    # the same without the fluid-velocity gather of the first sub-step of each
    # LB step, which is reported on its own (a grid operation, not per marker)
    gather_avg = ga_tot / max(args.steps, 1)
    ibm_wo = ibm_avg - ga_tot / max(ibm_n, 1)
    markers_per_s_wo = n / ((ibm_wo + spread_lb / md) * 1e-3) if n and ibm_wo > 0 else 0.0

    # same workload with the point particles moving: every MD sub-step
    # advances them by v * dt_MD (dt_MD = dt_LB / md_per_lb = 0.25; the
    # sphere has v = 0), on the context stream, and the sub-step sees new
    # positions: the markers are re-sorted, the interpolation and spreading
    # weights recomputed, the reactions of all sub-steps spread before the
    # next LB step, as for particles that a LAMMPS integrator moves
    ctx_stream = torch.cuda.ExternalStream(dev.stream())
    xm = tx.clone()

    def one_step_moving():
        dev.step(g, couplb_amd.RHO_REFRESHED_NEXT)
        for _ in range(md):
            with torch.cuda.stream(ctx_stream):
                xm.add_(tv, alpha=1.0 / md)
            dev.ibm_device(n, xm.data_ptr(), tv.data_ptr(), tm.data_ptr(), tf.data_ptr(),
                           positions_unchanged=False, **ibm_kw)

    mv_steps = max(8, args.steps // 2)
    dev.set_timer_mask(1 << 0)
    for _ in range(2):
        one_step_moving()
    dev.sync()
    t2 = time.perf_counter()
    for _ in range(mv_steps):
        one_step_moving()
    dev.sync()
    mlups_moving = nodes * mv_steps / (time.perf_counter() - t2) / 1e6

    # the fixed-marker steps with the marker arrays in host memory, as the
    # LAMMPS fix hands them over (it packs the group into pinned buffers):
    # every sub-step copies x, v, m to the device and the forces back, and
    # waits for them (PCIe-inclusive rate, not `value`)
    hx = torch.from_numpy(x).pin_memory().numpy()
    hv = torch.from_numpy(v).pin_memory().numpy()
    hm = torch.from_numpy(m).pin_memory().numpy()

    # the forces of every sub-step go into a cleared pinned buffer (LAMMPS
    # clears its forces itself before each force computation) and the fix is
    # told so (f is then not sent to the device); the masses do not change.
    # The buffers are cleared before the timed steps.
    host_steps = 24
    hfs = [torch.zeros((n, 3), dtype=torch.float64).pin_memory().numpy() for _ in range(md * (host_steps + 1))]

    def one_step_host(s, first=False):
        dev.step(g, couplb_amd.RHO_REFRESHED_NEXT)
        for k in range(md):
            dev.ibm(hx, hv, hm, hfs[s * md + k], positions_unchanged=not (first and k == 0),
                    masses_unchanged=True, forces_zero=True, **ibm_kw)

    # the first call follows the moving particles above: its positions are
    # new, and the device keeps the moving particles' marker ordering until
    # it is reorder_every (8) intervals old; the warm-up runs past that, so
    # that the timed steps hold the fixed-marker path (the warm-up's forces
    # go to the first buffers, which are written, not read: forces_zero)
    for w in range(10):
        one_step_host(0, first=w == 0)
    dev.sync()
    t3 = time.perf_counter()
    for st in range(1, host_steps + 1):
        one_step_host(st)
    dev.sync()
    mlups_host = nodes * host_steps / (time.perf_counter() - t3) / 1e6
    for b in hfs:
        b[:] = 0.0

    # moving particles with the marker arrays in host memory: the positions
    # of every MD sub-step (advanced by v * dt_MD as above) are prepared in
    # pinned buffers beforehand, so the timed steps hold the fix's own work
    # (uploads, sub-steps, force download), not the host's particle advance
    mh_steps = 24
    xs = [torch.from_numpy(x + v * ((k + 1) / md)).pin_memory().numpy() for k in range(md * (mh_steps + 1))]

    def one_step_moving_host(s):
        dev.step(g, couplb_amd.RHO_REFRESHED_NEXT)
        for k in range(md):
            dev.ibm(xs[s * md + k], hv, hm, hfs[s * md + k], positions_unchanged=False, masses_unchanged=True,
                    forces_zero=True, **ibm_kw)

    one_step_moving_host(0)
    dev.sync()
    t4 = time.perf_counter()
    for st in range(1, mh_steps + 1):
        one_step_moving_host(st)
    dev.sync()
    mlups_moving_host = nodes * mh_steps / (time.perf_counter() - t4) / 1e6
    del xs, hfs

    # achievable-bandwidth reference: a plain copy of one population buffer's
    # worth (k_step reads one state and writes the other)
    copy = dev.copy_bandwidth(nodes * Q * s_bytes, 10)

    # reduced storage: the error against the FP64 path on the same workload
    err = None
    if args.storage != "fp64":
        log("bench: error against the FP64 path")
        err = storage_error(N, args.storage, md, tx, tv, tm, ibm_kw, g)

    cpu = None
    if not args.no_cpu_baseline:
        log("bench: cpu baseline")
        cpu = cpu_baseline(N, args.cpu_nz, args.cpu_steps, md)
        cpu["host"] = host_cpu()
        # the same pass structure on one thread, on a thinner slab (its own
        # process: the OpenMP thread count is fixed when the library loads)
        import subprocess
        env = dict(os.environ, OMP_NUM_THREADS="1")
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-only", "8,10"], env=env,
                           capture_output=True, text=True, timeout=600)
        if r.returncode == 0 and r.stdout.strip():
            one = json.loads(r.stdout.strip().splitlines()[-1])
            cpu["one_thread"] = {"value": one["value"], "cores": one["cores"], "sample": one["sample"]}

    dtype = {"fp64": "f64", "fp32": "f32 storage/f32 math", "fp16": "f16 storage/f32 math",
             "bf16": "bf16 storage/f32 math", "i16": "16-bit fixed-point storage/f32 math"}[args.storage]
    out = {
        "metric": "D3Q19 MLUPS & %HBM peak at 256^3 w/ IBM; markers/s; err vs FP64 oracle",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic (seeded markers at prescribed fixed positions, uniform-flow initial state)",
        "config": {"workload": f"D3Q19 {N}^3 periodic + IBM: Fibonacci sphere D=24 ({nsphere} markers, "
                               f"fixed) + {args.particles} point particles held at prescribed fixed "
                               f"positions with constant Gaussian velocities, Peskin4, xi=1, "
                               f"md_per_lb={md}, u0=0.02, g_x=1e-6, nu=0.1 (moving particles: "
                               f"mlups_markers_moving)",
                   "grid": [N, N, N], "markers": n, "md_per_lb": md, "storage": args.storage},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": pmc_traffic(args.storage, N),
                     "kernel": f"k_step<3,{args.storage}> (fused pull + the fixed markers' force sums "
                               "+ collision)",
                     "bytes_per_lup": b_lup, "avg_kernel_ms": round(lb_avg, 4),
                     "device_copy_gbs": round(copy, 1), "frac_of_device_copy": round(achieved / copy, 4)},
        "ibm": {"markers_per_s": round(markers_per_s, 1), "avg_substep_ms": round(ibm_avg, 4),
                "avg_spread_ms_per_lb_step": round(spread_lb, 4),
                "spread_inside_lb_kernel_ms": round(spread_in_lb, 4), "markers": n,
                "gather_ms_per_lb_step": round(gather_avg, 4),
                "markers_per_s_without_gather": round(markers_per_s_wo, 1),
                "bytes_per_marker_substep": IBM_BYTES_PER_MARKER,
                "hbm_frac_without_gather": round(markers_per_s_wo * IBM_BYTES_PER_MARKER
                                                 / (HBM_PEAK_GBS * 1e9), 4)},
        "mlups_markers_moving": round(mlups_moving, 1),
        "mlups_marker_arrays_on_host": round(mlups_host, 1),
        "mlups_moving_marker_arrays_on_host": round(mlups_moving_host, 1),
        "lb_only_mlups": round(lb_only, 1),
        "lb_only_hbm_frac": round(lb_only * 1e6 * b_lup / (HBM_PEAK_GBS * 1e9), 4),
        "mass": diag["mass"],
        "err_vs_fp64": err,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
