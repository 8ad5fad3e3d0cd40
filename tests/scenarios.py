"""Shared test decks: one `fix couplb` command run by the oracle and by the
MI355X fix with identical atoms, in LAMMPS' Verlet order
(setup -> [f=0, post_force, end_of_step] per step)."""
from __future__ import annotations

import shlex

import numpy as np

import oracle


def parse_cmd(cmd):
    """Keyword subset of fix_couplb.cpp:106-234 -> oracle.Fix kwargs."""
    a = shlex.split(cmd)
    kw = dict(Nx=int(a[3]), Ny=int(a[4]), Nz=int(a[5]), nu=float(a[6]), rho=float(a[7]))
    walls = [[0, 0], [0, 0], [0, 0]]
    face = None
    i = 8
    while i < len(a):
        k = a[i]
        if k == "md_per_lb":
            kw["md_per_lb"] = int(a[i + 1]); i += 2
        elif k == "xi_ibm":
            kw["xi_ibm"] = float(a[i + 1]); i += 2
        elif k == "gravity":
            kw["gravity"] = tuple(float(x) for x in a[i + 1:i + 4]); i += 4
        elif k in ("wall_x", "wall_y", "wall_z"):
            walls["xyz".index(k[-1])] = [int(a[i + 1]), int(a[i + 2])]; i += 3
        elif k == "wall_vel":
            kw["wall_vel"] = tuple(float(x) for x in a[i + 1:i + 4]); i += 4
        elif k == "wall_vel_face":
            if face is None:
                face = [[[0.0] * 3, [0.0] * 3] for _ in range(3)]
            face["xyz".index(a[i + 1])][0 if a[i + 2] == "lo" else 1] = [float(x) for x in a[i + 3:i + 6]]
            i += 6
        elif k == "check_every":
            kw["check_every"] = int(a[i + 1]); i += 2
        elif k == "output":
            kw["output_every"] = int(a[i + 1]); i += 3
        elif k == "kernel":
            kw["kernel"] = oracle.ROMA if a[i + 1] == "roma" else oracle.PESKIN4; i += 2
        elif k in ("storage", "storage_range"):
            i += 2
        else:
            raise ValueError(k)
    kw["walls"] = tuple(tuple(w) for w in walls)
    if face is not None:
        kw["face_vel"] = face
    return kw


def atoms_from(x, v, m, mask=None):
    n = len(m)
    return dict(x=np.ascontiguousarray(x, np.float64), v=np.ascontiguousarray(v, np.float64),
                f=np.zeros((n, 3)), mask=np.ones(n, np.int32) if mask is None else mask.astype(np.int32),
                mass=np.ascontiguousarray(m, np.float64))


def copy_atoms(a):
    return {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in a.items()}


def run_oracle(cmd, steps, dimension=3, dt=1.0, atoms_fn=None, numerics="oracle", f0=None,
               boxlo=None, boxhi=None):
    kw = parse_cmd(cmd)
    fx = oracle.Fix(dim=dimension, dt=dt, numerics=numerics, boxlo=boxlo, boxhi=boxhi, **kw)
    if f0 is not None:
        fx.set_f(f0)
    a0 = atoms_fn(0) if atoms_fn else None
    fx.setup(a0)
    forces = []
    for _ in range(steps):
        fx.step += 1
        a = atoms_fn(fx.step) if atoms_fn else None
        if a is not None:
            a["f"][:] = 0.0
        fx.post_force(a)
        fx.end_of_step(fx.step)
        if a is not None:
            forces.append(a["f"].copy())
    return fx, forces


def run_device(cmd, steps, dimension=3, dt=1.0, atoms_fn=None, f0=None, boxlo=None, boxhi=None):
    import couplb_amd
    kw = {}
    if boxhi is not None:
        kw = dict(boxlo=boxlo, boxhi=boxhi)
    fx = couplb_amd.Fix(cmd, dimension=dimension, dt=dt, **kw)
    fx.init()
    if f0 is not None:
        import ctypes as C
        ctx = couplb_amd.lib().couplb_fix_context(fx.h)
        f = np.ascontiguousarray(f0, np.float64).reshape(-1)
        rc = couplb_amd.lib().couplb_set_state_f(ctx, f.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
    a0 = atoms_fn(0) if atoms_fn else None
    fx.setup(a0)
    forces = []
    for _ in range(steps):
        fx.step += 1
        a = atoms_fn(fx.step) if atoms_fn else None
        if a is not None:
            a["f"][:] = 0.0
        fx.post_force(a)
        fx.end_of_step(fx.step)
        if a is not None:
            forces.append(a["f"].copy())
    return fx, forces


def interior(f, dimension):
    """(Q, gz, gy, gx) -> interior populations."""
    if dimension == 3:
        return f[:, 1:-1, 1:-1, 1:-1]
    return f[:, :, 1:-1, 1:-1]


def rel_l2(a, b):
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


class DeviceArray:
    """Device buffer (FP64 unless dtype says otherwise) for the device-pointer
    entry points, allocated with
    the HIP runtime the library itself links (no second runtime in the
    process)."""
    _hip = None

    @classmethod
    def hip(cls):
        if cls._hip is None:
            import ctypes as C
            import couplb_amd
            couplb_amd.lib()            # loads libamdhip64 as a dependency
            h = C.CDLL("libamdhip64.so.7")
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipFree.argtypes = [C.c_void_p]
            cls._hip = h
        return cls._hip

    def __init__(self, a, dtype=np.float64):
        import ctypes as C
        self.host = np.ascontiguousarray(a, dtype)
        self.ptr = C.c_void_p()
        assert self.hip().hipMalloc(C.byref(self.ptr), max(8, self.host.nbytes)) == 0
        self.upload(self.host)

    def upload(self, a):
        a = np.ascontiguousarray(a, self.host.dtype)
        assert a.shape == self.host.shape
        assert self.hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0

    def download(self):
        out = np.empty_like(self.host)
        assert self.hip().hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, 2) == 0
        return out

    @property
    def addr(self):
        return self.ptr.value

    def __del__(self):
        if self._hip is not None and self.ptr:
            self._hip.hipFree(self.ptr)


def bad_density_state(n, seed=3):
    """Near-equilibrium populations with five broken nodes: all NaN, density
    -1, density 0 (next to the moving wall), one +Inf population and one
    -Inf population."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = n
    e, w = oracle.lattice(3)
    f = w[:, None, None, None] * (1.0 + 0.01 * rng.standard_normal((19, nz + 2, ny + 2, nx + 2)))
    f[:, 3, 3, 2] = np.nan
    f[:, 2, 4, 4] = -w
    f[:, 4, ny, 6] = 0.0
    f[1, 3, 2, 7] = np.inf
    f[3, 5, 5, 3] = -np.inf
    return f
