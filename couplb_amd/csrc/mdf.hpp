// mdf.hpp -- multi-direct-forcing IBM for markers with prescribed velocity.
//
// The reference couples once per sub-step with a penalty force
// (fix_couplb.cpp:779-785).  Multi-direct forcing instead finds the marker
// forces F_k (lattice units, force on the fluid) for which the velocity the
// next collision sees, u = (j + F/2)/rho (Guo, couplb_collision.h:28-33),
// interpolates to the prescribed marker velocity V_k:
//
//   r(F) = V - I(u* + S(F)/(2 rho)) = b - A F,   b = V - I(u*),
//   A F  = I0(S(F)/(2 rho))
//
// with u* = j/rho of the state, S = IBM::spread (couplb_ibm.h:92-133) over
// the sorted pairs of spread_pairs.hpp and I = IBM::interpolate
// (couplb_ibm.h:57-90) over four lanes per marker as in k_marker (I0: the
// same without the moving-wall velocity term, so A is linear).  A is
// symmetric positive semi-definite wherever the stencils see fluid only
// (interpolation and spreading then share the normalisation).  Two solvers:
//   Richardson  F += 2 omega r                 (classic multi-direct forcing;
//                                               one pass from F = 0 is the
//                                               reference's penalty force with
//                                               m xi/dt_LB = 2 omega)
//   CG          conjugate gradients on A F = b, restarted from the true
//               residual until max_k |r_k| < tol.
// The loops run on the device; a flag in device memory ends them, so the
// host queues several iterations between two convergence checks.
// Reductions are per block in a fixed order, then over blocks in a fixed
// order: results do not depend on scheduling.
#pragma once
#include "ibm.hpp"
#include "spread_pairs.hpp"

namespace couplb {

constexpr int MDF_THREADS = 256;

struct MdfScalars {
  double rr, pAp, alpha, beta, smax, smax_first;
  int done;    // 0 running, 1 converged, 2 breakdown (p.Ap <= 0), 3 iteration limit
  int iters;   // operator applications counted by the loop kernels
};

// block reduction of one value per thread: sum (op 0) or max (op 1), fixed
// order.  A butterfly inside each 64-lane wave (the two lanes of a step add
// the same two numbers, so every lane ends with the same bits), then the
// wave results in wave order.  The block has MDF_THREADS threads; sh holds
// one value per wave.
__device__ inline double block_reduce(double v, int op, double* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 64);
    v = op ? (v > w ? v : w) : v + w;
  }
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = sh[0];
#pragma unroll
  for (int i = 1; i < MDF_THREADS / 64; i++) r = op ? (r > sh[i] ? r : sh[i]) : r + sh[i];
  __syncthreads();
  return r;
}

// stencil geometry and weights of each marker (as k_marker, no velocity
// gather), the initial force F0 (caller order, lattice units; null = zero)
// into rec.f, Fm and P (the vector the operator is applied to), and the
// target velocity in sorted order, lattice units
template <int D>
__global__ void __launch_bounds__(256) k_mdf_init(int n, const int* perm, const double* x,
                                                  const double* V, const double* F0, IbmParams ip,
                                                  Geo g, Walls w, MarkerRec* rec, double* Vs,
                                                  double* Fm, double* P, double* gam) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int idx = perm[p];
  double lx, ly, lz;
  to_local(ip, D, x + 3 * (size_t)idx, lx, ly, lz);
  const bool fin = isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 &&
                   fabs(ly) < 1e9 && fabs(lz) < 1e9;
  MarkerRec r;
  r.i0 = fin ? (int)floor(lx) : -100;
  r.j0 = fin ? (int)floor(ly) : -100;
  r.k0 = (fin && D == 3) ? (int)floor(lz) : 0;
#pragma unroll
  for (int t = 0; t < 4; t++) {
    r.wx[t] = delta1(ip.kern, lx - (double)(r.i0 - 1 + t));
    r.wy[t] = delta1(ip.kern, ly - (double)(r.j0 - 1 + t));
    r.wz[t] = (D == 3) ? delta1(ip.kern, lz - (double)(r.k0 - 1 + t)) : 1.0;
  }
  // fluid weight sum per stencil plane, planes combined in order: the same
  // rounding as the four-lane sums of k_marker
  double wsf = 0.0, wsa = 0.0;  // fluid weight, weight of every counted node
  if (fin) {
#pragma unroll
    for (int v = 0; v < ((D == 3) ? 4 : 1); v++) {
      const int c = (D == 3) ? r.k0 - 1 + v : 0;
      double pl = 0.0, pla = 0.0;
      if (!(D == 3 && (c < 0 || c > g.nz + 1))) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const int b = r.j0 - 1 + u;
          if (b < 0 || b > g.ny + 1) continue;
#pragma unroll
          for (int t = 0; t < 4; t++) {
            const int a = r.i0 - 1 + t;
            if (a < 0 || a > g.nx + 1) continue;
            int wd, wsd;
            const int nt = node_type(w, g, a, b, c, wd, wsd);
            const double d = (D == 3) ? (r.wx[t] * r.wy[u]) * r.wz[v] : r.wx[t] * r.wy[u];
            if (nt == 0) pl += d;
            if (nt >= 0 && nt <= 3) pla += d;  // IBM::interpolate counts types 0-3
          }
        }
      }
      wsf += pl;
      wsa += pla;
    }
  }
  r.valid = (fin && !(wsf < DELTA_TOL)) ? 1 : 0;
  r.sc = r.valid ? 1.0 / wsf : 0.0;
  r.fx = (r.valid && F0) ? F0[3 * (size_t)idx + 0] : 0.0;
  r.fy = (r.valid && F0) ? F0[3 * (size_t)idx + 1] : 0.0;
  r.fz = (r.valid && F0) ? F0[3 * (size_t)idx + 2] : 0.0;
  rec[p] = r;
  gam[p] = r.valid ? wsf / wsa : 0.0;  // turns spreading weights into interpolation weights
  Fm[3 * (size_t)p + 0] = P[3 * (size_t)p + 0] = r.fx;
  Fm[3 * (size_t)p + 1] = P[3 * (size_t)p + 1] = r.fy;
  Fm[3 * (size_t)p + 2] = P[3 * (size_t)p + 2] = r.fz;
#pragma unroll
  for (int c = 0; c < 3; c++) Vs[3 * (size_t)p + c] = V[3 * (size_t)idx + c] / ip.vel_scale;
}

// node values of the operator, ut = [u*] + S(P)/(2 rho), one thread per
// node run of the sorted pairs
__global__ void __launch_bounds__(256) k_mdf_ut(int nh, const int* heads, const int* hlen,
                                                const unsigned* hnode, const int* pp,
                                                const double* pw, const double* P, const double* u4,
                                                int affine, double* ut, const MdfScalars* S) {
  if (S->done) return;
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  if (P) run_sum(heads[h], hlen[h], pp, pw, P, ax, ay, az);  // P null: F = 0
  const unsigned k = hnode[h];
  const double4 q = *reinterpret_cast<const double4*>(u4 + 4 * (size_t)k);
  const double hf = 0.5 / q.w;
  ut[3 * (size_t)k + 0] = (affine ? q.x : 0.0) + ax * hf;
  ut[3 * (size_t)k + 1] = (affine ? q.y : 0.0) + ay * hf;
  ut[3 * (size_t)k + 2] = (affine ? q.z : 0.0) + az * hf;
}

// Interpolate ut at every marker.  affine: out = r = Vs - I(ut), block
// partials (sum r.r, max |r|).  linear: out = A p = I0(ut), block partial
// sum p.Ap.
template <int D>
__global__ void __launch_bounds__(MDF_THREADS) k_mdf_interp(int n, Geo g, Walls w, const double* ut,
                                                            int affine, const MarkerRec* rec,
                                                            const double* Vs, const double* P,
                                                            double* out,
                                                            double* part_a, double* part_b,
                                                            const MdfScalars* S) {
  __shared__ double sh[MDF_THREADS / 64];
  if (S->done) return;
  constexpr int LPM = (D == 3) ? 4 : 1;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = tid / LPM, vv = tid % LPM;
  const bool live = p < n;
  const bool valid = live && rec[p].valid;
  double wx[4], wy[4], wzv = 0.0;
  int i0 = 0, j0 = 0, k0 = 0;
#pragma unroll
  for (int t = 0; t < 4; t++) { wx[t] = 0.0; wy[t] = 0.0; }
  if (valid) {
    const MarkerRec& r = rec[p];
#pragma unroll
    for (int t = 0; t < 4; t++) { wx[t] = r.wx[t]; wy[t] = r.wy[t]; }
    wzv = (D == 3) ? r.wz[vv] : 1.0;
    i0 = r.i0; j0 = r.j0; k0 = r.k0;
  }
  double s[5];
  interp_plane<D>(g, w, valid, i0, j0, k0, vv, wx, wy, wzv,
                  [&](unsigned node, int, double* un) {
                    un[0] = ut[3 * (size_t)node + 0];
                    un[1] = ut[3 * (size_t)node + 1];
                    un[2] = ut[3 * (size_t)node + 2];
                  }, s, affine != 0);
  combine_planes<LPM>(s);
  double a = 0.0, b = 0.0;
  if (live && vv == 0) {
    double o0 = 0.0, o1 = 0.0, o2 = 0.0;
    if (valid) {
      const double iw = 1.0 / s[3];
      const double u0 = s[0] * iw, u1 = s[1] * iw, u2 = s[2] * iw;
      if (affine) {
        o0 = Vs[3 * (size_t)p + 0] - u0;
        o1 = Vs[3 * (size_t)p + 1] - u1;
        o2 = Vs[3 * (size_t)p + 2] - u2;
        a = o0 * o0 + o1 * o1 + o2 * o2;
        b = sqrt(a);
      } else {
        o0 = u0; o1 = u1; o2 = u2;
        a = P[3 * (size_t)p + 0] * o0 + P[3 * (size_t)p + 1] * o1 + P[3 * (size_t)p + 2] * o2;
      }
    }
    out[3 * (size_t)p + 0] = o0;
    out[3 * (size_t)p + 1] = o1;
    out[3 * (size_t)p + 2] = o2;
  }
  a = block_reduce(a, 0, sh);
  b = block_reduce(b, 1, sh);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = a;
    part_b[blockIdx.x] = b;
  }
}

// Partials folded in a fixed order: thread t sums entries t, t + 256, ...,
// then the block tree of block_reduce.  Every block that folds the same
// partials finds the same value.
// (eight loads in flight per thread)
__device__ inline double fold_sum(const double* part, int nblk, double* sh) {
  double a = 0.0;
  for (int i0 = threadIdx.x; i0 < nblk; i0 += 8 * MDF_THREADS) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u * MDF_THREADS;
      v[u] = i < nblk ? part[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) a += v[u];
  }
  return block_reduce(a, 0, sh);
}
__device__ inline double fold_max(const double* part, int nblk, double* sh) {
  double b = 0.0;
  for (int i0 = threadIdx.x; i0 < nblk; i0 += 8 * MDF_THREADS) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int i = i0 + u * MDF_THREADS;
      v[u] = i < nblk ? part[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) b = b > v[u] ? b : v[u];
  }
  return block_reduce(b, 1, sh);
}

// One block: the true residual's partials into the scalars: rr, smax (the
// first time also smax_first); done = 1 when smax < tol, else 3 once maxit
// operator applications are counted.  respect_done: when done is already
// set (inside the Richardson loop), only copy the scalars.  mirror: copy of the
// scalars in pinned host memory, read by the host after a stream
// synchronisation without a device-to-host copy.
__global__ void __launch_bounds__(MDF_THREADS) k_mdf_scalar(const double* part_a,
                                                            const double* part_b, int nblk,
                                                            double tol, int first, int maxit,
                                                            int respect_done, MdfScalars* S,
                                                            MdfScalars* mirror) {
  __shared__ double sh[MDF_THREADS / 64];
  if (respect_done && S->done) {
    if (threadIdx.x == 0 && mirror) *mirror = *S;
    return;
  }
  const double a = fold_sum(part_a, nblk, sh);
  const double b = fold_max(part_b, nblk, sh);
  if (threadIdx.x != 0) return;
  S->rr = a;
  S->smax = b;
  if (first) S->smax_first = b;
  S->done = (b < tol) ? 1 : 0;
  S->iters += 1;
  if (!S->done && S->iters >= maxit) S->done = 3;
  if (mirror) *mirror = *S;
}

// Conjugate-gradient pass in three kernels: the operator (Ap = A p and the
// block partials of p.Ap), k_cg_update and k_cg_next.  A pass reads the
// scalars of its start, Sk, and k_cg_next writes the whole state after it
// into Sn; passes alternate between two buffers, so no block reads a field
// that another block of the same kernel writes.  Every block folds the
// partials it needs itself (fold_sum / fold_max), so all blocks agree on
// alpha and beta.  Markers without a valid stencil hold r = p = Ap = 0 and
// stay unchanged.

// alpha = rr / p.Ap; F += alpha p, r -= alpha Ap; block partials of r.r and
// max |r|.  p.Ap <= 0 (breakdown) leaves F and r and is seen by k_cg_next.
__global__ void __launch_bounds__(MDF_THREADS) k_cg_update(int n, const double* P, const double* Ap,
                                                           double* Fm, double* R, const double* px,
                                                           int npx, double* part_a, double* part_b,
                                                           MdfScalars* Sk) {
  __shared__ double sh[MDF_THREADS / 64];
  if (Sk->done) return;
  const double rr = Sk->rr;
  const double pAp = fold_sum(px, npx, sh);
  const bool ok = pAp > 0.0;
  const double al = ok ? rr / pAp : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Sk->pAp = pAp;  // fields no block of this kernel reads
    Sk->alpha = al;
  }
  if (!ok) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  double a = 0.0, b = 0.0;
  if (p < n) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      Fm[3 * (size_t)p + c] += al * P[3 * (size_t)p + c];
      const double rc = R[3 * (size_t)p + c] - al * Ap[3 * (size_t)p + c];
      R[3 * (size_t)p + c] = rc;
      a += rc * rc;
    }
    b = sqrt(a);
  }
  a = block_reduce(a, 0, sh);
  b = block_reduce(b, 1, sh);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = a;
    part_b[blockIdx.x] = b;
  }
}

// rr' and smax from k_cg_update's partials; done = 2 after a breakdown, 1
// when smax < tol, 3 at maxit applications; beta = rr'/rr and p = r + beta p
// while running.  Block 0 writes the new state to Sn (and the mirror).
__global__ void __launch_bounds__(MDF_THREADS) k_cg_next(int n, const double* R, double* P,
                                                         const double* part_a, const double* part_b,
                                                         int nblk, double tol, int maxit,
                                                         const MdfScalars* Sk, MdfScalars* Sn,
                                                         MdfScalars* mirror) {
  __shared__ double sh[MDF_THREADS / 64];
  MdfScalars s = *Sk;
  const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
  if (!s.done && !(s.pAp > 0.0)) s.done = 2;
  if (!s.done) {
    const double rr = fold_sum(part_a, nblk, sh);
    const double smax = fold_max(part_b, nblk, sh);
    s.beta = rr / s.rr;
    s.rr = rr;
    s.smax = smax;
    s.iters += 1;
    if (smax < tol) s.done = 1;
    else if (s.iters >= maxit) s.done = 3;
  }
  if (lead) {
    *Sn = s;
    if (mirror) *mirror = s;
  }
  if (s.done) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
#pragma unroll
  for (int c = 0; c < 3; c++) P[3 * (size_t)p + c] = R[3 * (size_t)p + c] + s.beta * P[3 * (size_t)p + c];
}

// restart: p = r
__global__ void __launch_bounds__(256) k_cg_restart(int n, const double* R, double* P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 3 * n) P[i] = R[i];
}

// Richardson step F += 2 omega r (after a true-residual evaluation that did
// not converge); P carries F
__global__ void __launch_bounds__(256) k_mdf_relax(int n, const double* R, double omega, double* Fm,
                                                   const MarkerRec* rec, double* P, const MdfScalars* S) {
  if (S->done) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n || !rec[p].valid) return;
  const double k2 = 2.0 * omega;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double f = Fm[3 * (size_t)p + c] + k2 * R[3 * (size_t)p + c];
    Fm[3 * (size_t)p + c] = f;
    P[3 * (size_t)p + c] = f;
  }
}

// P = F (before a true residual) and rec.f = F (the LB step's spreading)
__global__ void __launch_bounds__(256) k_mdf_setf(int n, const double* Fm, MarkerRec* rec, double* P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  MarkerRec& r = rec[p];
  double f[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    f[c] = r.valid ? Fm[3 * (size_t)p + c] : 0.0;
    P[3 * (size_t)p + c] = f[c];
  }
  r.fx = f[0]; r.fy = f[1]; r.fz = f[2];
}

__global__ void k_mdf_reset(MdfScalars* S, int clear_iters) {
  S->done = 0;
  if (clear_iters) S->iters = 0;
}

// force on the body per marker, caller order and units: -F * force_scale
__global__ void __launch_bounds__(256) k_mdf_out(int n, const int* perm, const double* Fm,
                                                 const MarkerRec* rec, double force_scale,
                                                 double* f_body) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int idx = perm[p];
  const bool v = rec[p].valid;
#pragma unroll
  for (int c = 0; c < 3; c++)
    f_body[3 * (size_t)idx + c] = v ? -Fm[3 * (size_t)p + c] * force_scale : 0.0;
}

}  // namespace couplb

namespace couplb {

// Inputs of the warm-start network for every marker (sorted order), from the
// slip at F = 0, s = V - I(u*) (R), sigma = |s|:
//   x[l]      = ((V - u_l) . s) / sigma^2   for the 64 stencil slots
//               l = t + 4u + 16v -> node (i0-1+t, j0-1+u, k0-1+v); u_l is u*
//               at fluid nodes, the wall velocity at moving-wall nodes, 0 at
//               other wall nodes; slots beyond the ghost layer (and the 48
//               unused slots in 2D) are 0
//   x[64..66] = s / sigma
// Markers that are not valid or have sigma < 1e-14 get zero inputs.
constexpr int MLP_IN = 67;
constexpr double MLP_FSCALE = 16.0;  // network output = F / (16 sigma)

template <int D>
__global__ void __launch_bounds__(256) k_mlp_features(int n, Geo g, Walls w, const MarkerRec* rec,
                                                      const double* u4, const double* Vs,
                                                      const double* R, float* X, double* sig) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  float* xr = X + (size_t)p * MLP_IN;
  const MarkerRec& r = rec[p];
  const double s0 = R[3 * (size_t)p + 0], s1 = R[3 * (size_t)p + 1], s2 = R[3 * (size_t)p + 2];
  const double ss = s0 * s0 + s1 * s1 + s2 * s2;
  const double sg = sqrt(ss);
  sig[p] = sg;
  if (!r.valid || !(sg >= 1e-14)) {
    for (int l = 0; l < MLP_IN; l++) xr[l] = 0.0f;
    return;
  }
  const double v0 = Vs[3 * (size_t)p + 0], v1 = Vs[3 * (size_t)p + 1], v2 = Vs[3 * (size_t)p + 2];
  const double inv = 1.0 / ss;
  for (int l = 0; l < 64; l++) {
    const int t = l & 3, u = (l >> 2) & 3, vv = l >> 4;
    float val = 0.0f;
    const int a = r.i0 - 1 + t, b = r.j0 - 1 + u, c = (D == 3) ? r.k0 - 1 + vv : 0;
    const bool in = (D == 3 || vv == 0) && a >= 0 && a <= g.nx + 1 && b >= 0 && b <= g.ny + 1 &&
                    (D == 2 || (c >= 0 && c <= g.nz + 1));
    if (in) {
      int wd, wsd;
      const int nt = node_type(w, g, a, b, c, wd, wsd);
      double u0 = 0.0, u1 = 0.0, u2 = 0.0;
      if (nt == 0) {
        const unsigned node = (unsigned)wrap_ghost(a, g.nx) +
                              (unsigned)g.nx * ((unsigned)wrap_ghost(b, g.ny) +
                                                (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0));
        const double4 q = *reinterpret_cast<const double4*>(u4 + 4 * (size_t)node);
        u0 = q.x; u1 = q.y; u2 = q.z;
      } else if (nt == 2) {
        u0 = w.vel[wd][wsd][0]; u1 = w.vel[wd][wsd][1]; u2 = (D == 3) ? w.vel[wd][wsd][2] : 0.0;
      }
      val = (float)(((v0 - u0) * s0 + (v1 - u1) * s1 + (v2 - u2) * s2) * inv);
    }
    xr[l] = val;
  }
  xr[64] = (float)(s0 / sg);
  xr[65] = (float)(s1 / sg);
  xr[66] = (float)(s2 / sg);
}

// F0 = 16 sigma y from the network's outputs (sorted order) into Fm, P and rec.f
__global__ void __launch_bounds__(256) k_mdf_warm(int n, const float* Y, const double* sig,
                                                  double* Fm, MarkerRec* rec, double* P) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  MarkerRec& r = rec[p];
  const bool ok = r.valid && sig[p] >= 1e-14;
  const double k = MLP_FSCALE * sig[p];
  double f[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    f[c] = ok ? k * (double)Y[3 * (size_t)p + c] : 0.0;
    Fm[3 * (size_t)p + c] = f[c];
    P[3 * (size_t)p + c] = f[c];
  }
  r.fx = f[0]; r.fy = f[1]; r.fz = f[2];
}

// network inputs (sorted) -> caller order
__global__ void __launch_bounds__(256) k_mlp_unsort(int n, const int* perm, const float* Xs, float* X) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * MLP_IN) return;
  const int p = (int)(e / MLP_IN), l = (int)(e % MLP_IN);
  X[(size_t)perm[p] * MLP_IN + l] = Xs[e];
}

// training targets F / (16 sigma), caller order (0 where sigma is tiny)
__global__ void __launch_bounds__(256) k_mdf_targets(int n, const int* perm, const double* Fm,
                                                     const double* sig, const MarkerRec* rec,
                                                     float* T) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const bool ok = rec[p].valid && sig[p] >= 1e-14;
  const double k = ok ? 1.0 / (MLP_FSCALE * sig[p]) : 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) T[3 * (size_t)perm[p] + c] = (float)(Fm[3 * (size_t)p + c] * k);
}

}  // namespace couplb
