"""couplb_diagnostics_noalloc (no device allocation: its partial sums live in
the library's fixed per-GPU scratch array) against couplb_diagnostics on the
same states, bit for bit: a large grid, then a smaller one on the same GPU
(the shared array keeps the first call's contents), then the large one
again after an LB step."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(d):
    import couplb_amd
    out = np.zeros(6)
    rc = couplb_amd.lib().couplb_diagnostics_noalloc(d.h, out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == 0, couplb_amd.lib().couplb_last_error(d.h).decode()
    return out


def _with_alloc(d):
    import couplb_amd
    out = np.zeros(6)
    assert couplb_amd.lib().couplb_diagnostics(d.h, out.ctypes.data_as(C.POINTER(C.c_double))) == 0
    return out


def _perturbed(dim, n, seed):
    import oracle
    rng = np.random.default_rng(seed)
    e, w = oracle.lattice(dim)
    shape = (len(w), n[2] + 2, n[1] + 2, n[0] + 2) if dim == 3 else (len(w), 1, n[1] + 2, n[0] + 2)
    return w.reshape((-1, 1, 1, 1)) * (1.0 + 0.02 * rng.standard_normal(shape))


def test_noalloc_diagnostics_equal_the_allocating_ones():
    import couplb_amd
    big = couplb_amd.Device(3, 48, 40, 36, tau=0.7, walls=((0, 0), (1, 2), (0, 0)),
                            face_vel=[[[0.0] * 3, [0.0] * 3], [[0.0] * 3, [0.01, 0.0, 0.0]],
                                      [[0.0] * 3, [0.0] * 3]])
    small = couplb_amd.Device(2, 20, 12, 1, tau=0.9, storage=couplb_amd.FP32)
    big.set_state(_perturbed(3, (48, 40, 36), 1))
    small.set_state(_perturbed(2, (20, 12, 1), 2))
    for d, name in ((big, "48x40x36 fp64"), (small, "20x12 fp32"), (big, "48x40x36 fp64 after a step")):
        if "after" in name:
            d.step((1e-5, 0.0, 0.0), 0)
        print("case noalloc diagnostics", name, flush=True)
        a = _noalloc(d)
        b = _with_alloc(d)
        print(name, a)
        assert np.array_equal(a, b)
        assert a[0] > 0 and a[4] > 0
