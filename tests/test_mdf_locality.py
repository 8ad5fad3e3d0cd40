"""Why a marker-local network cannot shorten the multi-direct-forcing solve.

On one sphere of the C5 configuration (diameter 16, 804 Fibonacci markers,
Peskin4, rho = 1 and all stencil nodes fluid, where the operator of the
solve is A = I S / 2 with the interpolation and spreading weights of
couplb_ibm.h:57-133) the converged forces F* = A^-1 b for slips b = -I u of
smooth velocity fields u are compared with what predictors can recover:

- CG from F = 0 needs ~150 passes to max slip 1e-6 (A's condition number is
  ~7,600: markers 1 dx apart under a 4 dx kernel);
- a start within 10 % of F* (random error) needs about half the passes;
- the best linear map from a marker's own 64 stencil velocities to its F*,
  fitted marker by marker, is within ~10 %: F* is local, but the map
  depends on the marker's neighbourhood;
- the best such map shared by all markers (what one network reading a
  marker's stencil velocities can represent, up to its nonlinearity) misses
  by > 80 %: the large components of F* lie along directions that the
  interpolation barely sees and that the neighbourhood sets.

What does shorten it on the same sphere: the inverses of A's diagonal blocks
over 32 markers in Morton order of their base cells (the device's block-Jacobi
preconditioned CG) take about half the passes of CG, where point Jacobi takes
the same number (A's diagonal is constant).

CPU only (numpy and the oracle's delta function).
"""
import math

import numpy as np

import oracle

NF = 600          # smooth velocity fields (400 to fit, 200 to test)


def _sphere_operator():
    n = int(round(math.pi * 16.0 ** 2))
    i = np.arange(n) + 0.5
    z = 1.0 - 2.0 * i / n
    r = np.sqrt(1.0 - z * z)
    phi = i * math.pi * (3.0 - math.sqrt(5.0))
    X = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1) * 8.0 + np.array([20.3, 20.7, 20.1])
    i0 = np.floor(X).astype(int)
    idx, stn, wst = {}, np.zeros((n, 64), int), np.zeros((n, 64))
    for k in range(n):
        d = [[oracle.delta(oracle.PESKIN4, X[k, a] - (i0[k, a] - 1 + t)) for t in range(4)] for a in range(3)]
        s = 0
        for c in range(4):
            for b in range(4):
                for a in range(4):
                    key = (i0[k, 0] - 1 + a, i0[k, 1] - 1 + b, i0[k, 2] - 1 + c)
                    stn[k, s] = idx.setdefault(key, len(idx))
                    wst[k, s] = d[0][a] * d[1][b] * d[2][c]
                    s += 1
        wst[k] /= wst[k].sum()
    S = np.zeros((len(idx), n))
    for k in range(n):
        S[stn[k], k] = wst[k]
    nodes = np.array(sorted(idx, key=idx.get), float)
    return 0.5 * S.T @ S, S, stn, nodes


def _cg_passes(A, b, x0, tol):
    x, r, it = x0.copy(), b - A @ x0, 1
    p, rr = r.copy(), float(r @ r)
    while np.abs(r).max() >= tol and it < 5000:
        Ap = A @ p
        it += 1
        a = rr / float(p @ Ap)
        x += a * p
        r -= a * Ap
        rn = float(r @ r)
        p = r + (rn / rr) * p
        rr = rn
    return it


def test_converged_forces_need_the_neighbourhood():
    A, S, stn, nodes = _sphere_operator()
    n = A.shape[0]
    rng = np.random.default_rng(5)
    U = np.zeros((len(nodes), NF))
    for f in range(NF):
        for _ in range(12):
            kv = rng.normal(size=3)
            kv *= rng.uniform(0, 2 * np.pi / 8) / np.linalg.norm(kv)
            U[:, f] += rng.normal() * np.cos(nodes @ kv + rng.uniform(0, 2 * np.pi))
    B = -(S.T @ U)
    Fs = np.linalg.solve(A, B)
    tr, te = slice(0, 400), slice(400, NF)

    # passes from zero and from a start 10 % off F*
    b = B[:, 0] * (0.01 / np.abs(B[:, 0]).max())      # slips ~1e-2, tolerance 1e-6
    f = np.linalg.solve(A, b)
    zero = _cg_passes(A, b, np.zeros(n), 1e-6)
    near = _cg_passes(A, b, f * (1 + 0.1 * rng.standard_normal(n)), 1e-6)
    print(f"CG passes: from zero {zero}, from F* with 10 % error {near}")
    assert zero > 100 and near < 0.6 * zero

    # marker by marker: 64 stencil velocities -> F*_k
    errs = []
    for k in range(0, n, 23):
        c, *_ = np.linalg.lstsq(U[stn[k], tr].T, Fs[k, tr], rcond=None)
        p = U[stn[k], te].T @ c
        errs.append(np.linalg.norm(p - Fs[k, te]) / np.linalg.norm(Fs[k, te]))
    local = float(np.mean(errs))
    # one map for all markers
    Xtr = np.concatenate([U[stn[k], tr].T for k in range(n)])
    c, *_ = np.linalg.lstsq(Xtr, np.concatenate([Fs[k, tr] for k in range(n)]), rcond=None)
    Xte = np.concatenate([U[stn[k], te].T for k in range(n)])
    yte = np.concatenate([Fs[k, te] for k in range(n)])
    shared = float(np.linalg.norm(Xte @ c - yte) / np.linalg.norm(yte))
    print(f"best linear predictor from a marker's stencil: per marker {local:.3f}, shared {shared:.3f}")
    assert local < 0.2 and shared > 0.8


def _morton(c):
    # 21 bits of each coordinate interleaved, x lowest (as the device's block order)
    key = np.zeros(len(c), dtype=np.int64)
    for bit in range(21):
        for d in range(3):
            key |= ((c[:, d] >> bit) & 1).astype(np.int64) << (3 * bit + d)
    return key


def _pcg_passes(A, b, tol, apply_m):
    x = np.zeros_like(b)
    r = b.copy()
    z = apply_m(r)
    p, rz, it = z.copy(), float(r @ z), 1
    while np.abs(r).max() >= tol and it < 5000:
        Ap = A @ p
        it += 1
        a = rz / float(p @ Ap)
        x += a * p
        r -= a * Ap
        z = apply_m(r)
        rzn = float(r @ z)
        p = z + (rzn / rz) * p
        rz = rzn
    return it, x


def test_block_jacobi_over_morton_blocks_shortens_cg():
    # the preconditioner of the device's COUPLB_MDF_PCG: the inverses of the
    # diagonal blocks of A over 32 markers taken in Morton order of their
    # base cells.  Point Jacobi gains nothing (A's diagonal is constant on a
    # sphere of fluid nodes); the blocks cut the passes, and the iteration
    # still reaches A^-1 b.
    A, S, stn, nodes = _sphere_operator()
    n = A.shape[0]
    X = (S.T @ nodes) / S.sum(0)[:, None]            # first moment of Peskin4: the marker
    order = np.argsort(_morton(np.floor(X).astype(np.int64) + 2), kind="stable")
    blocks = [order[i:i + 32] for i in range(0, n, 32)]
    inv = [np.linalg.inv(A[np.ix_(bl, bl)]) for bl in blocks]

    def block(r):
        z = np.empty_like(r)
        for bl, m in zip(blocks, inv):
            z[bl] = m @ r[bl]
        return z

    diag = A.diagonal()
    assert np.ptp(diag) < 1e-12 * diag.max()
    rng = np.random.default_rng(3)
    nodes_u = np.zeros(len(nodes))
    for _ in range(12):
        kv = rng.normal(size=3)
        kv *= rng.uniform(0, 2 * np.pi / 8) / np.linalg.norm(kv)
        nodes_u += rng.normal() * np.cos(nodes @ kv + rng.uniform(0, 2 * np.pi))
    b = -(S.T @ nodes_u)
    b *= 0.01 / np.abs(b).max()
    f = np.linalg.solve(A, b)
    cg, _ = _pcg_passes(A, b, 1e-6, lambda r: r)
    jac, _ = _pcg_passes(A, b, 1e-6, lambda r: r / diag)
    bj, x = _pcg_passes(A, b, 1e-6, block)
    print(f"passes to max slip 1e-6: CG {cg}, Jacobi {jac}, block-Jacobi (32, Morton) {bj}")
    assert abs(jac - cg) <= 0.05 * cg
    assert bj <= 0.7 * cg
    assert np.abs(A @ x - b).max() < 1e-6
    assert np.linalg.norm(x - f) < 1e-2 * np.linalg.norm(f)
