"""The numpy reference of the warm-start network (oracle/mlp.py), pinned
before it is used to check the device kernels: its analytic gradient against
central finite differences of its own float64 loss, its GELU against the
closed form's known values, its BF16 rounding against exact cases."""
import numpy as np

from oracle import mlp


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = mlp.init_params(1)
    X = rng.normal(0, 1, (7, mlp.IN))
    T = rng.normal(0, 1, (7, mlp.OUT))
    _, g = mlp.loss_grad(P, X, T, bf=False)
    picks = np.concatenate([rng.choice(np.arange(lo, hi), 12, replace=False) for lo, hi in
                            [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2),
                             (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_B2, mlp.OFF_W3),
                             (mlp.OFF_W3, mlp.OFF_B3)]] + [np.arange(mlp.OFF_B3, mlp.NPARAM)])
    h = 1e-6
    for i in picks:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp.loss_grad(Pp, X, T, bf=False)[0] - mlp.loss_grad(Pm, X, T, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(fd), (i, fd, g[i])


def test_gelu_values_and_derivative():
    assert mlp.gelu(np.array(0.0)) == 0.0
    # GELU(1) = Phi(1) = 0.8413447460685429...
    assert abs(mlp.gelu(np.array(1.0)) - 0.8413447460685429) < 1e-15
    z = np.linspace(-4, 4, 41)
    fd = (mlp.gelu(z + 1e-6) - mlp.gelu(z - 1e-6)) / 2e-6
    assert np.max(np.abs(fd - mlp.gelu_d(z))) < 1e-8


def test_bf16_rounding_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7                         # BF16 spacing in [1, 2)
    assert mlp.bf16(np.array([one + ulp / 2]))[0] == one            # tie -> even
    assert mlp.bf16(np.array([one + 3 * ulp / 2]))[0] == one + 2 * ulp   # tie -> even
    assert mlp.bf16(np.array([one + ulp / 2 + 2 ** -20]))[0] == one + ulp
    assert mlp.bf16(np.array([-3.0]))[0] == -3.0
