// ibm.hpp -- immersed-boundary kernels: marker ordering by base row, fluid
// velocity at covered nodes, one-marker-per-lane interpolation + penalty
// force + spread record.  Spreading itself happens inside the fused step
// (kernels.hpp, k_step) as a deterministic gather over the sorted records.
//
// All arithmetic that decides stencil placement or weights runs with FP
// contraction off, so floor() picks the same base cell and the delta values
// round exactly as the reference's couplb_ibm.h:11-55 compiled without FMA.
#pragma once
#include "kernels.hpp"

namespace couplb {

// delta_roma / delta_peskin4, couplb_ibm.h:11-33
__device__ inline double delta_roma(double r) {
#pragma clang fp contract(off)
  const double ar = fabs(r);
  if (ar <= 0.5) {
    double arg = 1.0 - 3.0 * r * r;
    if (arg < 0.0) arg = 0.0;
    return (1.0 + sqrt(arg)) / 3.0;
  }
  if (ar <= 1.5) {
    double arg = -2.0 + 6.0 * ar - 3.0 * r * r;
    if (arg < 0.0) arg = 0.0;
    return (5.0 - 3.0 * ar - sqrt(arg)) / 6.0;
  }
  return 0.0;
}

__device__ inline double delta_peskin4(double r) {
#pragma clang fp contract(off)
  const double ar = fabs(r);
  if (ar <= 1.0) {
    double arg = 1.0 + 4.0 * ar - 4.0 * ar * ar;
    if (arg < 0.0) arg = 0.0;
    return (3.0 - 2.0 * ar + sqrt(arg)) / 8.0;
  }
  if (ar <= 2.0) {
    double arg = -7.0 + 12.0 * ar - 4.0 * ar * ar;
    if (arg < 0.0) arg = 0.0;
    return (5.0 - 2.0 * ar - sqrt(arg)) / 8.0;
  }
  return 0.0;
}

__device__ inline double delta1(int kern, double r) {
  return kern == 0 ? delta_roma(r) : delta_peskin4(r);
}

struct IbmParams {
  double lo[3];
  double inv_dx;
  double xi_over_dtLB;
  double vel_scale;
  double inv_N;
  double force_scale;
  int kern;
};

// IBM::phys_to_local (couplb_ibm.h:49-55), single rank (offset 0)
__device__ inline void to_local(const IbmParams& ip, int D, const double* x,
                                double& lx, double& ly, double& lz) {
#pragma clang fp contract(off)
  lx = (x[0] - ip.lo[0]) * ip.inv_dx + 1.0;
  ly = (x[1] - ip.lo[1]) * ip.inv_dx + 1.0;
  lz = (D == 3) ? ((x[2] - ip.lo[2]) * ip.inv_dx + 1.0) : 0.0;
}

struct Base { int i0, j0, k0, valid; };

// Row key of the marker's base cell; markers whose stencil misses the
// ghost-inclusive grid get the sentinel nrows and sort last.
__global__ void k_keys(int n, const double* x, IbmParams ip, Geo g, unsigned* keys, int* vals) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double lx, ly, lz;
  to_local(ip, g.D, x + 3 * (size_t)p, lx, ly, lz);
  unsigned key = (unsigned)g.nrows;
  if (isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 && fabs(ly) < 1e9 && fabs(lz) < 1e9) {
    const int i0 = (int)floor(lx), j0 = (int)floor(ly), k0 = (g.D == 3) ? (int)floor(lz) : 0;
    const bool in = i0 >= -2 && i0 <= g.nx + 2 && j0 >= -2 && j0 <= g.ny + 2 &&
                    (g.D == 2 || (k0 >= -2 && k0 <= g.nz + 2));
    if (in) key = (unsigned)((j0 + 2) + g.nry * ((g.D == 3) ? (k0 + 2) : 0));
  }
  keys[p] = key;
  vals[p] = p;
}

// rowstart[r] = first sorted position with key >= r, r = 0..nrows
__global__ void k_rowstart(const unsigned* skeys, int n, int nrows, int* rowstart) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nrows) return;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skeys[mid] < (unsigned)r) lo = mid + 1; else hi = mid;
  }
  rowstart[r] = lo;
}

__global__ void k_bases(int n, const int* perm, const double* x, IbmParams ip, Geo g, Base* bases) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double lx, ly, lz;
  to_local(ip, g.D, x + 3 * (size_t)perm[p], lx, ly, lz);
  Base b;
  const bool fin = isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 && fabs(ly) < 1e9 && fabs(lz) < 1e9;
  b.i0 = fin ? (int)floor(lx) : -100;
  b.j0 = fin ? (int)floor(ly) : -100;
  b.k0 = (fin && g.D == 3) ? (int)floor(lz) : 0;
  b.valid = fin ? 1 : 0;
  bases[p] = b;
}

struct BaseSlot {
  const int* rowstart;
  const Base* rec;
};

// Fluid velocity u = j/rho (compute_macroscopic(false), couplb_lattice.h:
// 155-202) at every interior node that an image of some marker stencil
// covers, once per state (stamp = state counter).
template <int D, typename S>
__global__ void __launch_bounds__(256) k_ugather(PullCtx pc, BaseSlot bs, double* ux, double* uy,
                                                 double* uz, int* stamp, int cycle) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const Geo& g = pc.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  int xa[3];
  const int na = images(i + 1, g.nx, pc.w.type[0][0], pc.w.type[0][1], xa);
  bool covered = false;
  visit_row<D>(bs, g, pc.w, j + 1, k + 1, [&](const Base& m, int, int) {
    for (int ia = 0; ia < na; ia++) {
      const int t = xa[ia] - m.i0 + 1;
      if (t >= 0 && t <= 3) covered = true;
    }
  });
  if (!covered) return;
  const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
  if (stamp[node] == cycle) return;
  T f[Q];
  pull_node<D, S>(pc, i, j, k, f);
  double A = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double fv = (double)f[q];
    A += fv;
    jx += L::ex(q) * fv;
    jy += L::ey(q) * fv;
    jz += L::ez(q) * fv;
  }
  double r = A + pc.base;
  if (r <= RHO_MIN) r = RHO_MIN;
  if (!isfinite(r)) r = RHO_MIN;
  const double inv_r = 1.0 / r;
  ux[node] = jx * inv_r;
  uy[node] = jy * inv_r;
  uz[node] = jz * inv_r;
  stamp[node] = cycle;
}

// One marker per lane (markers in base-row order).  IBM::interpolate
// (couplb_ibm.h:57-90) + penalty force (fix_couplb.cpp:779-785) + the data
// IBM::spread needs (couplb_ibm.h:92-133), stored as a MarkerRec.
// fmode 0: f_atom[idx] += F (device f); 1: f_atom[idx] = F (host adds).
template <int D>
__global__ void __launch_bounds__(256) k_marker(int n, const int* perm, const double* x,
                                                const double* v, const double* m, double* f_atom,
                                                int fmode, IbmParams ip, Geo g, Walls w,
                                                const double* ux, const double* uy,
                                                const double* uz, MarkerRec* rec) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int idx = perm[p];
  double lx, ly, lz;
  to_local(ip, D, x + 3 * (size_t)idx, lx, ly, lz);
  const bool fin = isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 && fabs(ly) < 1e9 && fabs(lz) < 1e9;
  MarkerRec r;
  const int i0 = fin ? (int)floor(lx) : -100;
  const int j0 = fin ? (int)floor(ly) : -100;
  const int k0 = (fin && D == 3) ? (int)floor(lz) : 0;
  for (int t = 0; t < 4; t++) {
    r.wx[t] = delta1(ip.kern, lx - (double)(i0 - 1 + t));
    r.wy[t] = delta1(ip.kern, ly - (double)(j0 - 1 + t));
    r.wz[t] = (D == 3) ? delta1(ip.kern, lz - (double)(k0 - 1 + t)) : 1.0;
  }
  double ui = 0, vi = 0, wi = 0, ws = 0, wsf = 0;
  const int vlo = (D == 3) ? 0 : 1, vhi = (D == 3) ? 3 : 1;  // 2D: single plane c = 0
  for (int vv = vlo; vv <= vhi; vv++) {
    const int c = (D == 3) ? k0 - 1 + vv : 0;
    if (D == 3 && (c < 0 || c > g.nz + 1)) continue;
    for (int u = 0; u < 4; u++) {
      const int b = j0 - 1 + u;
      if (b < 0 || b > g.ny + 1) continue;
      for (int t = 0; t < 4; t++) {
        const int a = i0 - 1 + t;
        if (a < 0 || a > g.nx + 1) continue;
        int wd, wsd;
        const int nt = node_type(w, g, a, b, c, wd, wsd);
        const double d = (D == 3) ? (r.wx[t] * r.wy[u]) * r.wz[vv] : r.wx[t] * r.wy[u];
        if (nt == 0) {
          const unsigned node = (unsigned)wrap_ghost(a, g.nx) +
                                (unsigned)g.nx * ((unsigned)wrap_ghost(b, g.ny) +
                                                  (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0));
          ui += ux[node] * d;
          vi += uy[node] * d;
          wi += uz[node] * d;
          ws += d;
          wsf += d;
        } else if (nt == 1 || nt == 3) {
          ws += d;
        } else if (nt == 2) {
          ui += w.vel[wd][wsd][0] * d;
          vi += w.vel[wd][wsd][1] * d;
          wi += ((D == 3) ? w.vel[wd][wsd][2] : 0.0) * d;
          ws += d;
        }
      }
    }
  }
  double uf0 = 0, uf1 = 0, uf2 = 0;
  if (ws > DELTA_TOL) {
    const double iw = 1.0 / ws;
    uf0 = ui * iw; uf1 = vi * iw; uf2 = wi * iw;
  }
  const double* vp = v + 3 * (size_t)idx;
  const double mm = m[idx];
  const double Fx = mm * ip.xi_over_dtLB * (uf0 * ip.vel_scale - vp[0]);
  const double Fy = mm * ip.xi_over_dtLB * (uf1 * ip.vel_scale - vp[1]);
  const double Fz = mm * ip.xi_over_dtLB * (uf2 * ip.vel_scale - vp[2]);
  double* fa = f_atom + 3 * (size_t)idx;
  if (fmode == 0) { fa[0] += Fx; fa[1] += Fy; fa[2] += Fz; }
  else { fa[0] = Fx; fa[1] = Fy; fa[2] = Fz; }
  r.fx = -Fx * ip.inv_N / ip.force_scale;
  r.fy = -Fy * ip.inv_N / ip.force_scale;
  r.fz = -Fz * ip.inv_N / ip.force_scale;
  r.i0 = i0; r.j0 = j0; r.k0 = k0;
  r.valid = (fin && !(wsf < DELTA_TOL)) ? 1 : 0;
  r.sc = r.valid ? 1.0 / wsf : 0.0;
  rec[p] = r;
}

}  // namespace couplb
