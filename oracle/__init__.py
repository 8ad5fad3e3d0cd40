"""ctypes bindings for the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package: it is the oracle for the HIP path, never the path.

  Oracle(ops="oracle")  drives the C restatement in libcouplb_oracle.so
  Oracle(ops="ref")     drives the same orchestration on the reference's own
                        numerics headers compiled into _ref/libcouplb_ref.so

Both use the reference's ghost-inclusive layout f[q*ntotal + n],
n = i + gx*(j + gy*k) (CoupLB/couplb_lattice.h:82-136).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_ORACLE = os.path.join(HERE, "libcouplb_oracle.so")
LIB_REF = os.path.join(HERE, "_ref", "libcouplb_ref.so")

ROMA, PESKIN4 = 0, 1
F_RHO, F_UX, F_UY, F_UZ, F_FX, F_FY, F_FZ, F_BCUX, F_BCUY, F_BCUZ, F_F, F_FBUF = range(12)
I_TYPE, I_WALLDIM = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class GridParams(C.Structure):
    _fields_ = [("D", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("tau", C.c_double), ("wall", (C.c_int * 2) * 3),
                ("face_vel", ((C.c_double * 3) * 2) * 3), ("rho_lb", C.c_double)]


class FixParams(C.Structure):
    _fields_ = [("dim", C.c_int), ("Nx", C.c_int), ("Ny", C.c_int), ("Nz", C.c_int),
                ("nu", C.c_double), ("rho", C.c_double), ("md_per_lb", C.c_int),
                ("xi_ibm", C.c_double), ("grav_on", C.c_int * 3), ("g", C.c_double * 3),
                ("wall", (C.c_int * 2) * 3), ("wall_vel", C.c_double * 3),
                ("use_face_vel", C.c_int), ("face_vel", ((C.c_double * 3) * 2) * 3),
                ("kernel", C.c_int), ("check_every", C.c_int), ("output_every", C.c_int),
                ("boxlo", C.c_double * 3), ("boxhi", C.c_double * 3), ("dt", C.c_double),
                ("ntimestep", C.c_long), ("output_file", C.c_char_p)]


class Atoms(C.Structure):
    _fields_ = [("nlocal", C.c_int), ("x", _dp), ("v", _dp), ("f", _dp),
                ("mask", _ip), ("groupbit", C.c_int), ("mass", _dp)]


class Ops(C.Structure):
    _fields_ = [("name", C.c_char_p), ("create", C.c_void_p), ("destroy", C.c_void_p),
                ("dfield", C.CFUNCTYPE(_dp, C.c_void_p, C.c_int)),
                ("ifield", C.CFUNCTYPE(_ip, C.c_void_p, C.c_int)),
                ("dims", C.CFUNCTYPE(None, C.c_void_p, _ip, C.POINTER(C.c_long), _ip)),
                ("compute_macroscopic", C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int)),
                ("collide", C.CFUNCTYPE(None, C.c_void_p)),
                ("exchange", C.CFUNCTYPE(None, C.c_void_p)),
                ("stream", C.CFUNCTYPE(None, C.c_void_p)),
                ("exchange_velocity", C.CFUNCTYPE(None, C.c_void_p)),
                ("exchange_forces", C.CFUNCTYPE(None, C.c_void_p)),
                ("clear_forces", C.CFUNCTYPE(None, C.c_void_p)),
                ("interpolate", C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                            _dp, C.c_double, C.c_int, _dp)),
                ("spread", C.CFUNCTYPE(None, C.c_void_p, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double,
                                       _dp, C.c_double, C.c_double, C.c_int)),
                ("max_velocity", C.CFUNCTYPE(C.c_double, C.c_void_p)),
                ("diagnostics", C.CFUNCTYPE(None, C.c_void_p, _dp)),
                ("delta", C.CFUNCTYPE(C.c_double, C.c_int, C.c_double))]


_libs: dict[str, C.CDLL] = {}


def build(ref: bool = False) -> None:
    """Compile the oracle (and, with ref=True, the reference-header build)."""
    targets = ["all"] + (["ref"] if ref else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def ref_available() -> bool:
    return os.path.exists(LIB_REF)


def _lib(kind: str) -> C.CDLL:
    if kind in _libs:
        return _libs[kind]
    if kind == "oracle":
        if not os.path.exists(LIB_ORACLE):
            build()
        lib = C.CDLL(LIB_ORACLE)
        lib.clo_oracle_ops.restype = C.POINTER(Ops)
        lib.clo_fix_create.restype = C.c_void_p
        lib.clo_fix_create.argtypes = [C.POINTER(Ops), C.POINTER(FixParams), C.c_char_p, C.c_int]
        for name in ("clo_fix_setup", "clo_fix_post_force"):
            getattr(lib, name).argtypes = [C.c_void_p, C.POINTER(Atoms)]
        lib.clo_fix_end_of_step.argtypes = [C.c_void_p, C.c_long]
        lib.clo_fix_end_of_step.restype = C.c_int
        lib.clo_fix_destroy.argtypes = [C.c_void_p]
        lib.clo_fix_grid.argtypes = [C.c_void_p]
        lib.clo_fix_grid.restype = C.c_void_p
        lib.clo_fix_scalar.argtypes = [C.c_void_p, C.c_char_p]
        lib.clo_fix_scalar.restype = C.c_double
        lib.clo_fix_last_diag.argtypes = [C.c_void_p, _dp]
        lib.clo_fix_last_profile.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.clo_fix_set_f.argtypes = [C.c_void_p, _dp]
        lib.clo_fix_set_gravity.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        lib.clo_delta_roma.restype = C.c_double
        lib.clo_delta_roma.argtypes = [C.c_double]
        lib.clo_delta_peskin4.restype = C.c_double
        lib.clo_delta_peskin4.argtypes = [C.c_double]
        lib.clo_time_lbm_steps.restype = C.c_double
        lib.clo_time_lbm_steps.argtypes = [C.POINTER(Ops), C.c_void_p, C.c_int,
                                           C.c_double, C.c_double, C.c_double]
        lib.clo_omp_threads.restype = C.c_int
        lib.clo_write_checkpoint.restype = C.c_int
        lib.clo_write_checkpoint.argtypes = [C.c_char_p, C.c_int, C.c_int, _ip, _ip, _ip,
                                             C.c_long, _dp, C.c_long]
        lib.clo_read_checkpoint.restype = C.c_int
        lib.clo_read_checkpoint.argtypes = [C.c_char_p, C.c_int, C.c_int, _ip, _ip, _ip,
                                            C.POINTER(C.c_long), _dp, C.c_long, C.c_char_p, C.c_int]
    elif kind == "ref":
        lib = C.CDLL(LIB_REF)
        lib.clr_ops.restype = C.POINTER(Ops)
    else:
        raise ValueError(kind)
    _libs[kind] = lib
    return lib


def ops(kind: str = "oracle"):
    if kind == "oracle":
        return _lib("oracle").clo_oracle_ops()
    return _lib("ref").clr_ops()


def delta(kern: int, r: float) -> float:
    lib = _lib("oracle")
    return lib.clo_delta_roma(r) if kern == ROMA else lib.clo_delta_peskin4(r)


def _p(a):
    return a.ctypes.data_as(_dp)


class Fix:
    """FixCoupLB orchestration (fix_couplb.cpp) on one rank.

    Parameters mirror `fix ID group couplb Nx Ny Nz nu rho0 [keywords]`.
    """

    def __init__(self, Nx, Ny, Nz, nu, rho, *, dim=3, md_per_lb=1, xi_ibm=1.0,
                 gravity=None, walls=((0, 0), (0, 0), (0, 0)), wall_vel=(0.0, 0.0, 0.0),
                 face_vel=None, kernel=ROMA, check_every=0, output_every=0,
                 boxlo=None, boxhi=None, dt=1.0, ntimestep=0, output_file=None,
                 numerics="oracle"):
        self.lib = _lib("oracle")
        self.ops_ptr = ops(numerics)
        p = FixParams()
        p.dim, p.Nx, p.Ny, p.Nz = dim, Nx, Ny, Nz
        p.nu, p.rho, p.md_per_lb, p.xi_ibm = nu, rho, md_per_lb, xi_ibm
        if gravity is not None:
            for d in range(3):
                p.grav_on[d] = 1
                p.g[d] = gravity[d]
        for d in range(3):
            for s in range(2):
                p.wall[d][s] = walls[d][s]
        for c in range(3):
            p.wall_vel[c] = wall_vel[c]
        if face_vel is not None:
            p.use_face_vel = 1
            for d in range(3):
                for s in range(2):
                    for c in range(3):
                        p.face_vel[d][s][c] = face_vel[d][s][c]
        p.kernel, p.check_every, p.output_every = kernel, check_every, output_every
        boxlo = boxlo if boxlo is not None else (0.0, 0.0, 0.0 if dim == 3 else -0.5)
        boxhi = boxhi if boxhi is not None else (float(Nx), float(Ny), float(Nz) if dim == 3 else 0.5)
        for d in range(3):
            p.boxlo[d], p.boxhi[d] = boxlo[d], boxhi[d]
        p.dt, p.ntimestep = dt, ntimestep
        p.output_file = output_file.encode() if output_file else None
        self.params = p
        err = C.create_string_buffer(256)
        h = self.lib.clo_fix_create(self.ops_ptr, C.byref(p), err, 256)
        if not h:
            raise ValueError(err.value.decode())
        self.h = h
        self.step = ntimestep
        self.dim, self.N = dim, (Nx, Ny, Nz)
        self.g = self.lib.clo_fix_grid(h)
        gd = (C.c_int * 3)()
        nt = C.c_long()
        q = C.c_int()
        self.ops_ptr.contents.dims(self.g, gd, C.byref(nt), C.byref(q))
        self.gdims = (gd[0], gd[1], gd[2])
        self.ntotal, self.Q = nt.value, q.value
        self._atoms = None

    def close(self):
        if self.h:
            self.lib.clo_fix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def scalar(self, name):
        return self.lib.clo_fix_scalar(self.h, name.encode())

    # -- atoms -----------------------------------------------------------
    def _atoms_struct(self, x, v, f, mask, mass, groupbit=1):
        a = Atoms()
        a.nlocal = len(mass)
        a.x, a.v, a.f = _p(x), _p(v), _p(f)
        a.mask = mask.ctypes.data_as(_ip)
        a.groupbit = groupbit
        a.mass = _p(mass)
        return a

    def setup(self, atoms=None):
        a = self._mk(atoms)
        self.lib.clo_fix_setup(self.h, C.byref(a))

    def _mk(self, atoms):
        if atoms is None:
            z = np.zeros((0, 3))
            atoms = dict(x=z, v=z, f=z.copy(), mask=np.zeros(0, np.int32), mass=np.zeros(0))
        self._keep = atoms
        return self._atoms_struct(atoms["x"], atoms["v"], atoms["f"], atoms["mask"],
                                  atoms["mass"], atoms.get("groupbit", 1))

    def post_force(self, atoms=None):
        a = self._mk(atoms)
        self.lib.clo_fix_post_force(self.h, C.byref(a))

    def end_of_step(self, step):
        return self.lib.clo_fix_end_of_step(self.h, step)

    def run(self, nsteps, atoms_fn=None):
        """LAMMPS Verlet order: f=0 -> post_force -> end_of_step, per step."""
        for _ in range(nsteps):
            self.step += 1
            atoms = atoms_fn(self.step) if atoms_fn else None
            if atoms is not None:
                atoms["f"][:] = 0.0
            self.post_force(atoms)
            self.end_of_step(self.step)

    # -- state -----------------------------------------------------------
    def field(self, which):
        ptr = self.ops_ptr.contents.dfield(self.g, which)
        n = self.Q * self.ntotal if which in (F_F, F_FBUF) else self.ntotal
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def ifield(self, which):
        ptr = self.ops_ptr.contents.ifield(self.g, which)
        return np.ctypeslib.as_array(ptr, shape=(self.ntotal,)).copy()

    def f(self):
        gx, gy, gz = self.gdims
        return self.field(F_F).reshape(self.Q, gz, gy, gx)

    def interior(self, arr):
        """Strip ghosts from a (..., gz, gy, gx) array."""
        if self.dim == 3:
            return arr[..., 1:-1, 1:-1, 1:-1]
        return arr[..., :, 1:-1, 1:-1]

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert f.size == self.Q * self.ntotal
        self.lib.clo_fix_set_f(self.h, _p(f))

    def set_gravity(self, g):
        self.lib.clo_fix_set_gravity(self.h, *g)

    def last_diag(self):
        out = np.zeros(6)
        self.lib.clo_fix_last_diag(self.h, _p(out))
        return dict(Ma=out[0], mass=out[1], mom=out[2:5].copy(), clamps=int(out[5]))

    def last_profile(self):
        ny = self.N[1]
        r, u, v = np.zeros(ny), np.zeros(ny), np.zeros(ny)
        n = self.lib.clo_fix_last_profile(self.h, _p(r), _p(u), _p(v))
        return (r, u, v) if n else None

    def time_lbm_steps(self, steps, g_lb=(0.0, 0.0, 0.0)):
        return self.lib.clo_time_lbm_steps(self.ops_ptr, self.g, steps, *g_lb)


def omp_threads() -> int:
    return _lib("oracle").clo_omp_threads()


def feq(D, rho, u):
    """Equilibrium populations (couplb_lattice.h:137-144) for arrays rho, u[3]."""
    e, w = lattice(D)
    eu = np.tensordot(e.astype(np.float64), u, axes=([1], [0]))
    uu = (u * u).sum(axis=0)
    cs2 = 1.0 / 3.0
    return w.reshape((-1,) + (1,) * rho.ndim) * rho * (1.0 + eu / cs2 + 0.5 * eu * eu / (cs2 * cs2) - 0.5 * uu / cs2)


def lattice(D):
    """Velocity set and weights, couplb_lattice.h:24-75."""
    if D == 2:
        e = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0],
                      [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0]])
        w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    else:
        e = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0],
                      [1, 0, 1], [-1, 0, -1], [1, 0, -1], [-1, 0, 1],
                      [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]])
        w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    return e, w


def write_checkpoint(fname, D, Q, dims, glob, step, f):
    lib = _lib("oracle")
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
    a = (C.c_int * 3)(*dims)
    o = (C.c_int * 3)(0, 0, 0)
    g = (C.c_int * 3)(*glob)
    return bool(lib.clo_write_checkpoint(fname.encode(), D, Q, a, o, g, step, _p(f), f.size))


def read_checkpoint(fname, D, Q, dims, glob, nf):
    lib = _lib("oracle")
    f = np.zeros(nf)
    a = (C.c_int * 3)(*dims)
    o = (C.c_int * 3)(0, 0, 0)
    g = (C.c_int * 3)(*glob)
    step = C.c_long()
    err = C.create_string_buffer(128)
    ok = lib.clo_read_checkpoint(fname.encode(), D, Q, a, o, g, C.byref(step), _p(f), nf, err, 128)
    if not ok:
        raise IOError(err.value.decode())
    return step.value, f


class Grid:
    """One numerics-table grid (couplb_lattice.h Grid<L>) driven directly
    through the ops table: create with walls and per-face velocities, set
    the populations, and call the reference's building blocks."""

    def __init__(self, D, n, tau=0.8, walls=((0, 0), (0, 0), (0, 0)), face_vel=None,
                 numerics="oracle"):
        self.ops = ops(numerics).contents
        p = GridParams()
        p.D, p.nx, p.ny, p.nz, p.tau, p.rho_lb = D, n[0], n[1], n[2] if D == 3 else 1, tau, 1.0
        for d in range(3):
            for s in range(2):
                p.wall[d][s] = walls[d][s]
                for c in range(3):
                    p.face_vel[d][s][c] = face_vel[d][s][c] if face_vel is not None else 0.0
        create = C.CFUNCTYPE(C.c_void_p, C.POINTER(GridParams))(self.ops.create)
        self.g = create(C.byref(p))
        self._destroy = C.CFUNCTYPE(None, C.c_void_p)(self.ops.destroy)
        gd = (C.c_int * 3)()
        nt = C.c_long()
        q = C.c_int()
        self.ops.dims(self.g, gd, C.byref(nt), C.byref(q))
        self.gdims, self.nt, self.Q, self.D = (gd[0], gd[1], gd[2]), nt.value, q.value, D

    def field(self, which):
        n = self.Q * self.nt if which in (F_F, F_FBUF) else self.nt
        return np.ctypeslib.as_array(self.ops.dfield(self.g, which), shape=(n,))

    def close(self):
        if getattr(self, "g", None):
            self._destroy(self.g)
            self.g = None

    def __del__(self):
        self.close()


def mdf_solve(grid, x, V, lo, inv_dx, kern, tol, relax=1.0, max_iter=200, vel_scale=1.0,
              F0=None, method="richardson"):
    """Multi-direct forcing restated over the reference's own interpolate /
    spread / exchange building blocks (couplb_ibm.h:57-133,
    couplb_streaming.h:140-158).  No reference counterpart: the reference
    couples with one penalty pass.  With u* = j/rho of the grid's state the
    slip of a marker force set F (force on the fluid, lattice units) is

      r(F) = V/vel_scale - I(u* + S(F)/(2 rho))

    richardson: F += 2 relax r until max_k |r_k| < tol (no update after the
    converged evaluation); cg: conjugate gradients on A F = r(0) with
    A p = I(u* + S(p)/(2 rho)) - I(u*), restarted from the true slip when
    the recursive one converges.  Iterations count slip / operator
    evaluations.  Returns (F, iterations, slip_first, slip_last); the grid
    is left with F spread into its force fields (folded) and u* in its
    velocity fields."""
    o, g = grid.ops, grid.g
    n = len(x)
    x = np.ascontiguousarray(x, np.float64)
    Vl = np.ascontiguousarray(V, np.float64) / vel_scale
    F = np.zeros((n, 3)) if F0 is None else np.array(F0, np.float64)
    lo_a = (C.c_double * 3)(*lo)
    out = (C.c_double * 3)()
    o.compute_macroscopic(g, 0)
    o.exchange_velocity(g)
    rho = grid.field(F_RHO).copy()
    ustar = [grid.field(k).copy() for k in (F_UX, F_UY, F_UZ)]
    fields_u = [grid.field(k) for k in (F_UX, F_UY, F_UZ)]
    fields_f = [grid.field(k) for k in (F_FX, F_FY, F_FZ)]
    dx = 1.0 / inv_dx
    dv = dx ** grid.D

    def spread(Fs):
        o.clear_forces(g)
        for k in range(n):
            o.spread(g, x[k, 0], x[k, 1], x[k, 2], Fs[k, 0], Fs[k, 1], Fs[k, 2], dv, lo_a, inv_dx,
                     dx, kern)
        o.exchange_forces(g)

    def interp_at(Fs):
        """I(u* + S(Fs)/(2 rho)) at every marker"""
        spread(Fs)
        for c in range(3):
            fields_u[c][:] = ustar[c] + fields_f[c] * (0.5 / rho)
        o.exchange_velocity(g)
        U = np.empty((n, 3))
        for k in range(n):
            o.interpolate(g, x[k, 0], x[k, 1], x[k, 2], lo_a, inv_dx, kern, out)
            U[k] = out[:]
        return U

    smax = lambda r: float(np.sqrt((r ** 2).sum(1)).max())
    U0 = interp_at(np.zeros((n, 3)))
    r = Vl - interp_at(F)
    it = 1
    slip_first = slip_last = smax(r)
    if method == "richardson":
        while slip_last >= tol and it < max_iter:
            F = F + 2.0 * relax * r
            r = Vl - interp_at(F)
            it += 1
            slip_last = smax(r)
    else:
        while slip_last >= tol and it < max_iter:
            p = r.copy()
            rr = float((r * r).sum())
            while it < max_iter:
                Ap = interp_at(p) - U0
                pAp = float((p * Ap).sum())
                if not pAp > 0:
                    break
                alpha = rr / pAp
                F = F + alpha * p
                r = r - alpha * Ap
                it += 1
                rr_new = float((r * r).sum())
                if smax(r) < tol:
                    break
                p = r + (rr_new / rr) * p
                rr = rr_new
            r = Vl - interp_at(F)     # true slip
            it += 1
            slip_last = smax(r)
    spread(F)
    for c in range(3):
        fields_u[c][:] = ustar[c]
    return F, it, slip_first, slip_last
