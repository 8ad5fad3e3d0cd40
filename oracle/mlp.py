"""Reference computation of the warm-start network (TEST INFRASTRUCTURE ONLY).

The network of include/couplb_mlp.h in float64 numpy: 67 -> 64 -> 64 -> 3,
GELU(z) = z/2 (1 + erf(z/sqrt 2)), loss = mean over samples of the squared
error summed over the outputs.  With bf16=True the values the device feeds
to its BF16 matrix-core products (inputs, weights, activations, back-
propagated errors) are rounded to BF16 (round to nearest even) at the same
points, so the device can be compared with it; with bf16=False it is the
plain float64 network whose gradient finite differences pin.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

IN, H, OUT = 67, 64, 3
OFF_W1 = 0
OFF_B1 = OFF_W1 + IN * H
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * OUT
NPARAM = OFF_B3 + OUT


def bf16(a):
    """Round to BF16 (nearest even) through float32, returned as float64."""
    f = np.ascontiguousarray(a, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def unpack(P):
    P = np.asarray(P, np.float64)
    return dict(W1=P[OFF_W1:OFF_B1].reshape(IN, H), b1=P[OFF_B1:OFF_W2],
                W2=P[OFF_W2:OFF_B2].reshape(H, H), b2=P[OFF_B2:OFF_W3],
                W3=P[OFF_W3:OFF_B3].reshape(H, OUT), b3=P[OFF_B3:NPARAM])


def forward(P, X, bf=True):
    q = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, np.float64))
    p = unpack(P)
    Xb, W1, W2, W3 = q(X), q(p["W1"]), q(p["W2"]), q(p["W3"])
    Z1 = Xb @ W1 + p["b1"]
    A1 = q(gelu(Z1))
    Z2 = A1 @ W2 + p["b2"]
    A2 = q(gelu(Z2))
    Y = A2 @ W3 + p["b3"]
    return dict(Xb=Xb, W1=W1, W2=W2, W3=W3, Z1=Z1, A1=A1, Z2=Z2, A2=A2, Y=Y)


def loss_grad(P, X, T, bf=True):
    """(loss, gradient as a parameter vector)"""
    q = (lambda a: bf16(a)) if bf else (lambda a: np.asarray(a, np.float64))
    f = forward(P, X, bf)
    n = len(X)
    d = f["Y"] - T
    loss = float((d * d).sum() / n)
    dYf = 2.0 * d / n
    dY = q(dYf)
    g = np.zeros(NPARAM)
    g[OFF_B3:NPARAM] = dYf.sum(0)
    g[OFF_W3:OFF_B3] = (f["A2"].T @ dY).reshape(-1)
    dZ2f = (dY @ f["W3"].T) * gelu_d(f["Z2"])
    dZ2 = q(dZ2f)
    g[OFF_B2:OFF_W3] = dZ2f.sum(0)
    g[OFF_W2:OFF_B2] = (f["A1"].T @ dZ2).reshape(-1)
    dZ1f = (dZ2 @ f["W2"].T) * gelu_d(f["Z1"])
    dZ1 = q(dZ1f)
    g[OFF_B1:OFF_W2] = dZ1f.sum(0)
    g[OFF_W1:OFF_B1] = (f["Xb"].T @ dZ1).reshape(-1)
    return loss, g


def init_params(seed):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM)
    P[OFF_W1:OFF_B1] = rng.normal(0, 1 / np.sqrt(IN), IN * H)
    P[OFF_W2:OFF_B2] = rng.normal(0, 1 / np.sqrt(H), H * H)
    P[OFF_W3:OFF_B3] = rng.normal(0, 1 / np.sqrt(H), H * OUT)
    P[OFF_B1:OFF_W2] = rng.normal(0, 0.1, H)
    P[OFF_B2:OFF_W3] = rng.normal(0, 0.1, H)
    P[OFF_B3:NPARAM] = rng.normal(0, 0.1, OUT)
    return P
