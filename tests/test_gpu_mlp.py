"""Warm-start network kernels (include/couplb_mlp.h) against the numpy
reference oracle/mlp.py, which rounds to BF16 where the device does and
whose own gradient is pinned by finite differences (test_mlp_oracle.py)."""
import numpy as np
import pytest

from oracle import mlp as ref
from tests.scenarios import DeviceArray, rel_l2

pytestmark = pytest.mark.gpu


def _net(P):
    import couplb_amd
    net = couplb_amd.Mlp(seed=0)
    net.params = P
    return net


def _data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, ref.IN)).astype(np.float32)
    T = rng.normal(0, 1, (n, ref.OUT)).astype(np.float32)
    return X, T


def test_forward_matches_reference():
    P = ref.init_params(3).astype(np.float32)
    X, _ = _data(300, 1)                 # not a multiple of the 64-sample tile
    net = _net(P)
    dX, dY = DeviceArray(X, np.float32), DeviceArray(np.zeros((300, 3)), np.float32)
    net.forward_device(300, dX.addr, dY.addr)
    Y = dY.download()
    Yr = ref.forward(P.astype(np.float64), X.astype(np.float64))["Y"]
    print("forward rel-L2", rel_l2(Y, Yr), "max", np.max(np.abs(Y - Yr)))
    assert rel_l2(Y, Yr) < 1e-4


def test_gradient_matches_reference():
    P = ref.init_params(4).astype(np.float32)
    X, T = _data(200, 2)
    net = _net(P)
    dX, dT = DeviceArray(X, np.float32), DeviceArray(T, np.float32)
    loss, g = net.gradient_device(200, dX.addr, dT.addr)
    lr_, gr = ref.loss_grad(P.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert loss == pytest.approx(lr_, rel=1e-5)
    for name, lo, hi in [("W1", ref.OFF_W1, ref.OFF_B1), ("b1", ref.OFF_B1, ref.OFF_W2),
                         ("W2", ref.OFF_W2, ref.OFF_B2), ("b2", ref.OFF_B2, ref.OFF_W3),
                         ("W3", ref.OFF_W3, ref.OFF_B3), ("b3", ref.OFF_B3, ref.NPARAM)]:
        e = rel_l2(g[lo:hi].astype(np.float64), gr[lo:hi])
        print(f"grad {name}: rel-L2 {e:.2e}")
        assert e < 1e-3, name
    # a plausible slip (a transposed block) would be O(1) off: check the
    # comparison can see it
    assert rel_l2(g[ref.OFF_W2:ref.OFF_B2].reshape(64, 64).T.reshape(-1).astype(np.float64),
                  gr[ref.OFF_W2:ref.OFF_B2]) > 0.1


def test_training_fits_a_teacher_and_is_deterministic():
    n = 16384
    rng = np.random.default_rng(5)
    X = rng.normal(0, 1, (n, ref.IN)).astype(np.float32)
    teacher = ref.init_params(9)
    T = ref.forward(teacher, X.astype(np.float64), bf=False)["Y"].astype(np.float32)
    dX, dT = DeviceArray(X, np.float32), DeviceArray(T, np.float32)
    P0 = ref.init_params(10).astype(np.float32)
    out = []
    for _ in range(2):
        net = _net(P0)
        l0, _ = net.gradient_device(n, dX.addr, dT.addr)
        l1 = net.train_device(n, dX.addr, dT.addr, batch=512, epochs=20, lr=2e-3, seed=7)
        lf, _ = net.gradient_device(n, dX.addr, dT.addr)
        out.append((l0, l1, lf, net.params))
    (l0, l1, lf, p1), (_, _, _, p2) = out
    print("loss initial", l0, "last epoch", l1, "final", lf)
    assert lf < 0.1 * l0
    assert np.array_equal(p1, p2)          # fixed-order reductions: bit-identical runs
