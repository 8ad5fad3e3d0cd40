"""couplb_amd -- MI355X-native LB/IBM step for the CoupLB LAMMPS fix.

Python bindings (ctypes) over the C ABI of couplb_amd/libcouplb_hip.so:

  Device  -- the device context of include/couplb_hip.h (grid, LB step,
             IBM sub-step, diagnostics, state import/export)
  Fix     -- the host-side FixCoupLB of include/couplb_fix.h, driven with the
             reference's own command syntax:
               Fix("flow all couplb 16 64 1 0.1 1.0 wall_y 1 1 ...", domain)

The library is required: importing works without it, but every call raises
if libcouplb_hip.so has not been built (python -m couplb_amd.build) -- there
is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import shlex

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcouplb_hip.so")

OK, EINVAL, EDEVICE, ENOMEM, ESTATE = 0, -1, -2, -3, -4
FP64, FP32, FP16, BF16, I16 = 0, 1, 2, 3, 4
ROMA, PESKIN4 = 0, 1
STORAGE = {"fp64": FP64, "fp32": FP32, "fp16": FP16, "bf16": BF16, "i16": I16}
RHO_REFRESHED_NEXT = 1

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
MLP_INPUTS, MLP_OUTPUTS, MLP_NPARAM = 67, 3, 8707


class Params(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("tau", C.c_double), ("wall", (C.c_int * 2) * 3),
                ("face_vel", ((C.c_double * 3) * 2) * 3), ("kernel", C.c_int),
                ("storage", C.c_int), ("rho0", C.c_double), ("md_per_lb", C.c_int),
                ("device", C.c_int), ("storage_range", C.c_double)]


class IbmArgs(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("inv_dx", C.c_double), ("xi_over_dtLB", C.c_double),
                ("vel_scale", C.c_double), ("inv_N", C.c_double), ("force_scale", C.c_double),
                ("positions_unchanged", C.c_int), ("masses_unchanged", C.c_int), ("forces_zero", C.c_int)]


MDF_RICHARDSON, MDF_CG, MDF_PCG = 0, 1, 2


class MdfArgs(C.Structure):
    _fields_ = [("ibm", IbmArgs), ("tol", C.c_double), ("relax", C.c_double),
                ("max_iter", C.c_int), ("method", C.c_int)]


class MdfStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("slip_first", C.c_double),
                ("slip_last", C.c_double), ("slip_zero", C.c_double)]


class MdfLearn(C.Structure):
    _fields_ = [("warm_start", C.c_void_p), ("features", C.c_void_p), ("targets", C.c_void_p)]


class Domain(C.Structure):
    _fields_ = [("dimension", C.c_int), ("boxlo", C.c_double * 3), ("boxhi", C.c_double * 3),
                ("periodicity", C.c_int * 3), ("dt", C.c_double), ("ntimestep", C.c_long)]


class Atoms(C.Structure):
    _fields_ = [("nlocal", C.c_int), ("x", _dp), ("v", _dp), ("f", _dp),
                ("mask", C.POINTER(C.c_int)), ("groupbit", C.c_int), ("rmass", _dp),
                ("mass", _dp), ("type", C.POINTER(C.c_int))]


class AtomExtras(C.Structure):
    _fields_ = [("tag", C.c_void_p), ("tag_bytes", C.c_int), ("molecule", C.c_void_p),
                ("radius", _dp), ("mu", _dp), ("omega", _dp), ("torque", _dp)]


VAR_EVAL = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_char_p, _dp)

_lib = None


class CoupLBError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libcouplb_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise CoupLBError(f"{LIB_PATH} not built: run `python -m couplb_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, ip, dp = C.c_void_p, C.c_int, _dp
    sig = {
        "couplb_create": (ip, [C.POINTER(Params), C.POINTER(vp)]),
        "couplb_destroy": (ip, [vp]),
        "couplb_last_error": (C.c_char_p, [vp]),
        "couplb_version": (C.c_char_p, []),
        "couplb_set_state_f": (ip, [vp, dp]),
        "couplb_get_state_f": (ip, [vp, dp]),
        "couplb_state_len": (C.c_long, [vp]),
        "couplb_init_equilibrium": (ip, [vp, C.c_double, C.c_double, C.c_double, C.c_double]),
        "couplb_lbm_step": (ip, [vp, dp, ip]),
        "couplb_lbm_steps_async": (ip, [vp, dp, ip, ip]),
        "couplb_lbm_fuses_steps": (ip, [vp]),
        "couplb_set_timer_mask": (ip, [vp, C.c_uint]),
        "couplb_host_alloc": (vp, [C.c_size_t]),
        "couplb_host_free": (None, [vp]),
        "couplb_copy_bandwidth": (ip, [vp, C.c_size_t, ip, dp]),
        "couplb_ibm_substep": (ip, [vp, ip, dp, dp, dp, dp, C.POINTER(IbmArgs)]),
        "couplb_ibm_substep_device": (ip, [vp, ip, vp, vp, vp, vp, C.POINTER(IbmArgs)]),
        "couplb_refresh_macro": (ip, [vp]),
        "couplb_get_macro": (ip, [vp, dp, dp, dp, dp]),
        "couplb_get_node_forces": (ip, [vp, dp, ip]),
        "couplb_diagnostics": (ip, [vp, dp]),
        "couplb_diagnostics_noalloc": (ip, [vp, dp]),
        "couplb_profile_y": (ip, [vp, dp, dp, dp]),
        "couplb_synchronize": (ip, [vp]),
        "couplb_storage_saturations": (ip, [vp, C.POINTER(C.c_ulonglong)]),
        "couplb_stream": (vp, [vp]),
        "couplb_last_kernel_ms": (C.c_double, [vp, ip]),
        "couplb_kernel_ms_total": (ip, [vp, ip, dp, C.POINTER(C.c_long)]),
        "couplb_kernel_ms_reset": (ip, [vp]),
        "couplb_ibm_mdf_device": (ip, [vp, ip, vp, vp, vp, vp, C.POINTER(MdfArgs),
                                       C.POINTER(MdfLearn), C.POINTER(MdfStats)]),
        "couplb_mlp_create": (ip, [ip, C.c_ulonglong, C.POINTER(vp)]),
        "couplb_mlp_destroy": (ip, [vp]),
        "couplb_mlp_last_error": (C.c_char_p, [vp]),
        "couplb_mlp_get_params": (ip, [vp, _fp]),
        "couplb_mlp_set_params": (ip, [vp, _fp]),
        "couplb_mlp_forward_device": (ip, [vp, ip, vp, vp]),
        "couplb_mlp_gradient_device": (ip, [vp, ip, vp, vp, _fp, dp]),
        "couplb_mlp_train_device": (ip, [vp, ip, vp, vp, ip, ip, C.c_float, C.c_ulonglong, dp]),
        "couplb_fix_create": (ip, [ip, C.POINTER(C.c_char_p), C.POINTER(Domain), C.POINTER(vp)]),
        "couplb_fix_destroy": (ip, [vp]),
        "couplb_fix_last_error": (C.c_char_p, [vp]),
        "couplb_fix_set_variable_callback": (ip, [vp, VAR_EVAL, vp]),
        "couplb_fix_init": (ip, [vp]),
        "couplb_fix_set_domain": (ip, [vp, C.POINTER(Domain)]),
        "couplb_fix_setup": (ip, [vp, C.POINTER(Atoms)]),
        "couplb_fix_post_force": (ip, [vp, C.POINTER(Atoms)]),
        "couplb_fix_end_of_step": (ip, [vp, C.c_long]),
        "couplb_fix_set_atom_extras": (ip, [vp, C.POINTER(AtomExtras)]),
        "couplb_fix_setmask": (ip, []),
        "couplb_fix_scalar": (C.c_double, [vp, C.c_char_p]),
        "couplb_fix_last_diag": (ip, [vp, dp]),
        "couplb_fix_last_profile": (ip, [vp, dp, dp, dp]),
        "couplb_fix_context": (vp, [vp]),
        "couplb_fix_log": (C.c_char_p, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype, f.argtypes = res, args
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


def lattice_q(dim):
    return 19 if dim == 3 else 9


class Device:
    """One grid on one GPU (include/couplb_hip.h)."""

    def __init__(self, dim, nx, ny, nz, tau, walls=((0, 0), (0, 0), (0, 0)),
                 face_vel=None, kernel=ROMA, storage=FP64, rho0=1.0, md_per_lb=1, device=0,
                 storage_range=0.0):
        L = lib()
        p = Params()
        p.dim, p.nx, p.ny, p.nz, p.tau = dim, nx, ny, nz if dim == 3 else 1, tau
        for d in range(3):
            for s in range(2):
                p.wall[d][s] = walls[d][s]
                for c in range(3):
                    p.face_vel[d][s][c] = face_vel[d][s][c] if face_vel is not None else 0.0
        p.kernel = kernel
        p.storage = STORAGE[storage] if isinstance(storage, str) else storage
        p.rho0, p.md_per_lb, p.device = rho0, md_per_lb, device
        p.storage_range = storage_range
        h = C.c_void_p()
        rc = L.couplb_create(C.byref(p), C.byref(h))
        if rc != OK:
            msg = L.couplb_last_error(h).decode() if h.value else "couplb_create failed"
            if h.value:
                L.couplb_destroy(h)
            raise CoupLBError(f"couplb_create: {msg} (rc={rc})")
        self.h = h
        self.dim, self.n = dim, (nx, ny, nz if dim == 3 else 1)
        self.Q = lattice_q(dim)
        self.gdims = (nx + 2, ny + 2, (nz + 2) if dim == 3 else 1)
        self.params = p

    def _chk(self, rc, what):
        if rc != OK:
            raise CoupLBError(f"{what}: {lib().couplb_last_error(self.h).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "h", None):
            lib().couplb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state in the reference layout (Q, gz, gy, gx)
    def set_state(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1)
        assert f.size == lib().couplb_state_len(self.h)
        self._chk(lib().couplb_set_state_f(self.h, _p(f)), "set_state_f")

    def get_state(self):
        gx, gy, gz = self.gdims
        out = np.empty(self.Q * gx * gy * gz)
        self._chk(lib().couplb_get_state_f(self.h, _p(out)), "get_state_f")
        return out.reshape(self.Q, gz, gy, gx)

    def init_equilibrium(self, rho=1.0, u=(0.0, 0.0, 0.0)):
        self._chk(lib().couplb_init_equilibrium(self.h, rho, *u), "init_equilibrium")

    def step(self, g=(0.0, 0.0, 0.0), flags=0):
        ga = np.asarray(g, dtype=np.float64)
        self._chk(lib().couplb_lbm_step(self.h, _p(ga), flags), "lbm_step")

    def steps_async(self, n, g=(0.0, 0.0, 0.0), flags=0):
        ga = np.asarray(g, dtype=np.float64)
        self._chk(lib().couplb_lbm_steps_async(self.h, _p(ga), flags, n), "lbm_steps_async")

    def fuses_steps(self):
        """True when steps_async runs the steps after the first in one launch."""
        return bool(lib().couplb_lbm_fuses_steps(self.h))

    def sync(self):
        self._chk(lib().couplb_synchronize(self.h), "synchronize")

    def copy_bandwidth(self, nbytes=2 << 30, reps=10):
        """Read + write GB/s of a plain device copy of nbytes (measurement helper)."""
        out = np.zeros(1)
        self._chk(lib().couplb_copy_bandwidth(self.h, nbytes, reps, _p(out)), "copy_bandwidth")
        return float(out[0])

    def ibm(self, x, v, m, f, lo, inv_dx, xi_over_dtLB, vel_scale, inv_N, force_scale,
            positions_unchanged=False, masses_unchanged=False, forces_zero=False):
        """Host-array IBM sub-step; f ([n,3] float64) is added into in place.
        masses_unchanged / forces_zero: m equals the previous host sub-step's,
        f holds zeros (neither is then sent to the device)."""
        a = IbmArgs()
        for d in range(3):
            a.lo[d] = lo[d]
        a.inv_dx, a.xi_over_dtLB, a.vel_scale = inv_dx, xi_over_dtLB, vel_scale
        a.inv_N, a.force_scale, a.positions_unchanged = inv_N, force_scale, int(positions_unchanged)
        a.masses_unchanged, a.forces_zero = int(masses_unchanged), int(forces_zero)
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        m = np.ascontiguousarray(m, np.float64)
        assert f.dtype == np.float64 and f.flags.c_contiguous
        self._chk(lib().couplb_ibm_substep(self.h, len(m), _p(x), _p(v), _p(m), _p(f), C.byref(a)),
                  "ibm_substep")

    def ibm_device(self, n, x_ptr, v_ptr, m_ptr, f_ptr, lo, inv_dx, xi_over_dtLB, vel_scale,
                   inv_N, force_scale, positions_unchanged=False):
        """Device-pointer IBM sub-step (arrays resident in HBM)."""
        a = IbmArgs()
        for d in range(3):
            a.lo[d] = lo[d]
        a.inv_dx, a.xi_over_dtLB, a.vel_scale = inv_dx, xi_over_dtLB, vel_scale
        a.inv_N, a.force_scale, a.positions_unchanged = inv_N, force_scale, int(positions_unchanged)
        self._chk(lib().couplb_ibm_substep_device(self.h, n, x_ptr, v_ptr, m_ptr, f_ptr, C.byref(a)),
                  "ibm_substep_device")

    def mdf_device(self, n, x_ptr, V_ptr, F0_ptr, fbody_ptr, lo, inv_dx, vel_scale, force_scale,
                   tol=1e-6, relax=1.0, max_iter=200, positions_unchanged=False, method="cg",
                   warm_start=None, features_ptr=None, targets_ptr=None):
        """Multi-direct-forcing solve for markers with prescribed velocity V
        (device pointers; F0_ptr may be None); method "cg", "pcg" (block-Jacobi
        preconditioned) or "richardson";
        warm_start an Mlp; features_ptr / targets_ptr optional float32 device
        outputs for training.  Returns the solver statistics."""
        a = MdfArgs()
        for d in range(3):
            a.ibm.lo[d] = lo[d]
        a.ibm.inv_dx, a.ibm.vel_scale, a.ibm.force_scale = inv_dx, vel_scale, force_scale
        a.ibm.xi_over_dtLB, a.ibm.inv_N = 0.0, 1.0
        a.ibm.positions_unchanged = int(positions_unchanged)
        a.tol, a.relax, a.max_iter = tol, relax, max_iter
        methods = {"richardson": MDF_RICHARDSON, "cg": MDF_CG, "pcg": MDF_PCG}
        if method not in methods:
            raise ValueError(f"unknown MDF method {method!r}: richardson, cg or pcg")
        a.method = methods[method]
        st = MdfStats()
        learn = None
        if warm_start is not None or features_ptr or targets_ptr:
            learn = MdfLearn(warm_start.h.value if warm_start is not None else None,
                             features_ptr, targets_ptr)
        self._chk(lib().couplb_ibm_mdf_device(self.h, n, x_ptr, V_ptr, F0_ptr, fbody_ptr,
                                              C.byref(a), C.byref(learn) if learn else None,
                                              C.byref(st)), "ibm_mdf_device")
        return dict(iterations=st.iterations, converged=bool(st.converged),
                    slip_first=st.slip_first, slip_last=st.slip_last, slip_zero=st.slip_zero)

    def storage_saturations(self):
        n = C.c_ulonglong()
        self._chk(lib().couplb_storage_saturations(self.h, C.byref(n)), "storage_saturations")
        return n.value

    def refresh(self):
        self._chk(lib().couplb_refresh_macro(self.h), "refresh_macro")

    def macro(self):
        nx, ny, nz = self.n
        out = [np.empty(nx * ny * nz) for _ in range(4)]
        self._chk(lib().couplb_get_macro(self.h, *[_p(o) for o in out]), "get_macro")
        return [o.reshape(nz, ny, nx) for o in out]

    def node_forces(self, fold=True):
        """Node forces of the IBM records since the last LB step, [nz, ny, nx, 3]
        (fold: periodic ghost images added into the interior)."""
        nx, ny, nz = self.n
        out = np.empty(3 * nx * ny * nz)
        self._chk(lib().couplb_get_node_forces(self.h, _p(out), int(bool(fold))), "get_node_forces")
        return out.reshape(nz, ny, nx, 3)

    def diagnostics(self):
        out = np.zeros(6)
        self._chk(lib().couplb_diagnostics(self.h, _p(out)), "diagnostics")
        return dict(mass=out[0], mom=out[1:4].copy(), umax=out[4], clamps=int(out[5]))

    def profile(self):
        ny = self.n[1]
        r, u, v = np.zeros(ny), np.zeros(ny), np.zeros(ny)
        self._chk(lib().couplb_profile_y(self.h, _p(r), _p(u), _p(v)), "profile_y")
        return r, u, v

    def stream(self):
        return lib().couplb_stream(self.h)

    def last_ms(self, which=0):
        return lib().couplb_last_kernel_ms(self.h, which)

    def kernel_ms_total(self, which=0):
        """(total milliseconds, count) of the timed regions of one kind since
        the last reset: 0 LB kernel, 1 IBM sub-step, 2 spreading, 3 MDF, 4 fluid
        velocity gather at covered nodes (inside 1 or 3)."""
        t, n = C.c_double(), C.c_long()
        self._chk(lib().couplb_kernel_ms_total(self.h, which, C.byref(t), C.byref(n)),
                  "kernel_ms_total")
        return t.value, n.value

    def set_timer_mask(self, mask):
        """Kinds of kernel_ms_total that record events (bit k for kind k)."""
        self._chk(lib().couplb_set_timer_mask(self.h, mask), "set_timer_mask")

    def kernel_ms_reset(self):
        self._chk(lib().couplb_kernel_ms_reset(self.h), "kernel_ms_reset")


class Mlp:
    """Warm-start network of the multi-direct-forcing IBM (include/couplb_mlp.h).
    Array arguments of the *_device methods are device pointers."""

    def __init__(self, seed=0, device=0):
        h = C.c_void_p()
        rc = lib().couplb_mlp_create(device, seed, C.byref(h))
        self.h = h
        if rc != OK:
            raise CoupLBError(f"mlp_create: {lib().couplb_mlp_last_error(h).decode()} (rc={rc})")

    def _chk(self, rc, what):
        if rc != OK:
            raise CoupLBError(f"{what}: {lib().couplb_mlp_last_error(self.h).decode()} (rc={rc})")

    def close(self):
        if getattr(self, "h", None):
            lib().couplb_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def params(self):
        p = np.zeros(MLP_NPARAM, np.float32)
        self._chk(lib().couplb_mlp_get_params(self.h, p.ctypes.data_as(_fp)), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.shape == (MLP_NPARAM,)
        self._chk(lib().couplb_mlp_set_params(self.h, p.ctypes.data_as(_fp)), "set_params")

    def forward_device(self, n, X_ptr, Y_ptr):
        self._chk(lib().couplb_mlp_forward_device(self.h, n, X_ptr, Y_ptr), "forward")

    def gradient_device(self, n, X_ptr, T_ptr):
        g = np.zeros(MLP_NPARAM, np.float32)
        loss = C.c_double()
        self._chk(lib().couplb_mlp_gradient_device(self.h, n, X_ptr, T_ptr, g.ctypes.data_as(_fp),
                                                   C.byref(loss)), "gradient")
        return loss.value, g

    def train_device(self, n, X_ptr, T_ptr, batch=4096, epochs=1, lr=1e-3, seed=0):
        loss = C.c_double()
        self._chk(lib().couplb_mlp_train_device(self.h, n, X_ptr, T_ptr, batch, epochs, lr, seed,
                                                C.byref(loss)), "train")
        return loss.value


class Fix:
    """FixCoupLB host core (include/couplb_fix.h) driven like a LAMMPS deck.

    cmd: "ID group couplb Nx Ny Nz nu rho [keywords]" (the `fix` line
    without the word `fix`).  domain: dict(dimension, boxlo, boxhi,
    periodicity, dt, ntimestep).
    """

    def __init__(self, cmd, dimension=3, boxlo=(0, 0, 0), boxhi=None, periodicity=(1, 1, 1),
                 dt=1.0, ntimestep=0, variables=None):
        L = lib()
        args = shlex.split(cmd) if isinstance(cmd, str) else list(cmd)
        self._argv = (C.c_char_p * len(args))(*[a.encode() for a in args])
        d = Domain()
        d.dimension = dimension
        if boxhi is None and len(args) >= 6:
            N = [float(args[3]), float(args[4]), float(args[5])]
            boxhi = (N[0], N[1], N[2] if dimension == 3 else 0.5)
            if dimension == 2:
                boxlo = (boxlo[0], boxlo[1], -0.5)
        if boxhi is None:
            boxhi = (1.0, 1.0, 1.0)
        for k in range(3):
            d.boxlo[k], d.boxhi[k], d.periodicity[k] = boxlo[k], boxhi[k], periodicity[k]
        d.dt, d.ntimestep = dt, ntimestep
        self.domain = d
        self.step = ntimestep
        self.dimension = dimension
        self.Ny = int(args[4]) if len(args) > 4 else 0
        h = C.c_void_p()
        rc = L.couplb_fix_create(len(args), self._argv, C.byref(d), C.byref(h))
        self.h = h
        if rc != OK:
            msg = L.couplb_fix_last_error(h).decode()
            L.couplb_fix_destroy(h)
            self.h = None
            raise CoupLBError(msg)
        self.variables = dict(variables or {})

        def _eval(user, name, out):
            key = name.decode()
            if key not in self.variables:
                return 1
            if not out:         # init: existence check only
                return 0
            val = self.variables[key]
            out[0] = float(val(self.step) if callable(val) else val)
            return 0

        self._cb = VAR_EVAL(_eval)
        L.couplb_fix_set_variable_callback(h, self._cb, None)

    def _chk(self, rc):
        if rc != OK:
            raise CoupLBError(lib().couplb_fix_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().couplb_fix_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self):
        self._chk(lib().couplb_fix_init(self.h))

    @staticmethod
    def _atoms(atoms):
        a = Atoms()
        if atoms is None:
            a.nlocal = 0
            z = np.zeros((1, 3))
            a.x = a.v = a.f = _p(z)
            a._keep = (z,)
            a.mask = (C.c_int * 1)(0)
            a.groupbit = 1
            a.rmass = _p(np.ones(1))
            return a
        a.nlocal = len(atoms["mass"])
        a.x, a.v, a.f = _p(atoms["x"]), _p(atoms["v"]), _p(atoms["f"])
        a.mask = atoms["mask"].ctypes.data_as(C.POINTER(C.c_int))
        a.groupbit = atoms.get("groupbit", 1)
        a.rmass = _p(atoms["mass"])
        if "type" in atoms:
            a.type = atoms["type"].ctypes.data_as(C.POINTER(C.c_int))
        return a

    def set_atom_extras(self, tag=None, molecule=None, radius=None, mu=None, omega=None, torque=None):
        """Per-atom arrays for vtk_solid (int32 tag / molecule; float64
        radius [n], mu [n, 4], omega and torque [n, 3]); kept alive here."""
        e = AtomExtras()
        keep = []
        for name in ("tag", "molecule"):
            arr = locals()[name]
            if arr is not None:
                arr = np.ascontiguousarray(arr, dtype=np.int32)
                keep.append(arr)
                setattr(e, name, arr.ctypes.data)
        e.tag_bytes = 4
        for name in ("radius", "mu", "omega", "torque"):
            arr = locals()[name]
            if arr is not None:
                arr = np.ascontiguousarray(arr, dtype=np.float64)
                keep.append(arr)
                setattr(e, name, _p(arr))
        self._extras_keep = keep
        self._chk(lib().couplb_fix_set_atom_extras(self.h, C.byref(e)))

    def setup(self, atoms=None):
        self._keep = atoms
        self._chk(lib().couplb_fix_setup(self.h, C.byref(self._atoms(atoms))))

    def post_force(self, atoms=None):
        self._keep = atoms
        self._chk(lib().couplb_fix_post_force(self.h, C.byref(self._atoms(atoms))))

    def end_of_step(self, step):
        self._chk(lib().couplb_fix_end_of_step(self.h, step))

    def run(self, nsteps, atoms_fn=None):
        """LAMMPS Verlet order per step: f = 0 -> post_force -> end_of_step."""
        for _ in range(nsteps):
            self.step += 1
            atoms = atoms_fn(self.step) if atoms_fn else None
            if atoms is not None:
                atoms["f"][:] = 0.0
            self.post_force(atoms)
            self.end_of_step(self.step)

    def scalar(self, name):
        return lib().couplb_fix_scalar(self.h, name.encode())

    def last_diag(self):
        out = np.zeros(6)
        self._chk(lib().couplb_fix_last_diag(self.h, _p(out)))
        return dict(Ma=out[0], mass=out[1], mom=out[2:5].copy(), clamps=int(out[5]))

    def last_profile(self):
        r, u, v = np.zeros(self.Ny), np.zeros(self.Ny), np.zeros(self.Ny)
        rc = lib().couplb_fix_last_profile(self.h, _p(r), _p(u), _p(v))
        return (r, u, v) if rc == OK else None

    def log(self):
        return lib().couplb_fix_log(self.h).decode()

    def device_state(self):
        """Post-stream populations in the reference layout (Q, gz, gy, gx)."""
        ctx = lib().couplb_fix_context(self.h)
        if not ctx:
            raise CoupLBError("fix not initialised")
        n = lib().couplb_state_len(ctx)
        out = np.empty(n)
        rc = lib().couplb_get_state_f(ctx, _p(out))
        if rc != OK:
            raise CoupLBError(lib().couplb_last_error(ctx).decode())
        return out


def version() -> str:
    return lib().couplb_version().decode()
