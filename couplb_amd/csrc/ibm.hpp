// ibm.hpp -- immersed-boundary kernels: marker ordering by base row, fluid
// velocity at covered nodes, four-lanes-per-marker interpolation + penalty
// force + spread record.  Spreading is a deterministic gather: over the
// node-sorted (marker, stencil node) map inside the LB step (k_step, from
// spread_pairs.hpp's maps) and, for records off the map, over binned
// records per tile of nodes (spread_rows.hpp).
//
// All arithmetic that decides stencil placement or weights runs with FP
// contraction off, so floor() picks the same base cell and the delta values
// round exactly as the reference's couplb_ibm.h:11-55 compiled without FMA.
#pragma once
#include "spread_rows.hpp"

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

// Which interior nodes some marker stencil covers: one thread per (marker,
// stencil node) sets cover = 1 at the interior node that a fluid or
// periodic-image stencil node (node type 0, as IBM::interpolate reads it,
// couplb_ibm.h:66-86) stands for; the caller has cleared the array.  Every
// write stores 1, so overlapping stencils leave the same bytes in any order.
template <int D>
__global__ void __launch_bounds__(256) k_cover_scatter(int n, Geo g, Walls w, const Base* bases,
                                                       unsigned char* cover) {
  constexpr int S = stencil_size<D>();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * S) return;
  const int p = (int)(e / S), s = (int)(e % S);
  const Base b = bases[p];
  if (!b.valid) return;
  const int a = b.i0 - 1 + (s & 3), bb = b.j0 - 1 + ((s >> 2) & 3);
  const int c = (D == 3) ? b.k0 - 1 + (s >> 4) : 0;
  const bool in = a >= 0 && a <= g.nx + 1 && bb >= 0 && bb <= g.ny + 1 && (D == 2 || (c >= 0 && c <= g.nz + 1));
  int wd, wsd;
  if (in && node_type(w, g, a, bb, c, wd, wsd) == 0)
    cover[(unsigned)wrap_ghost(a, g.nx) +
          (unsigned)g.nx * ((unsigned)wrap_ghost(bb, g.ny) +
                            (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0))] = 1;
}

// Fluid velocity u = j/rho and density (compute_macroscopic(false),
// couplb_lattice.h:155-202) at covered nodes, once per state (stamp = state
// counter).  Stored node-interleaved (ux, uy, uz, rho): a stencil row is one
// contiguous segment for the marker gathers.
// DENSE: every lane pulls (full-row, fully coalesced reads) and only covered
// nodes are written; for coverage high enough that nearly every cache line
// of the populations is needed anyway.  Otherwise only covered nodes not yet
// computed for this state (stamp) pull.
// u = j/rho and rho from the pulled populations (compute_macroscopic(false))
template <int D, typename S>
__device__ inline double4 upull_moments(const typename Store<S>::T (&f)[Lat<D>::Q], double base) {
  using L = Lat<D>;
  constexpr int Q = L::Q;
  double A = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double fv = (double)f[q];
    A += fv;  // c.f with the nonzero terms only: the same sums, fewer FP64 ops
    if (L::ex(q) > 0) jx += fv; else if (L::ex(q) < 0) jx -= fv;
    if (L::ey(q) > 0) jy += fv; else if (L::ey(q) < 0) jy -= fv;
    if (L::ez(q) > 0) jz += fv; else if (L::ez(q) < 0) jz -= fv;
  }
  if (!isfinite(A)) moments_full<D>(f, jx, jy, jz);  // 0 * Inf = NaN, as the reference
  double r = A + base;
  if (r <= RHO_MIN) r = RHO_MIN;
  if (!isfinite(r)) r = RHO_MIN;
  const double inv_r = 1.0 / r;
  double4 out;
  out.x = jx * inv_r;
  out.y = jy * inv_r;
  out.z = jz * inv_r;
  out.w = r;
  return out;
}

// fresh: first gather of this state, every stamp is older (no stamp reads)
template <int D, typename S, bool DENSE>
__global__ void __launch_bounds__(256) k_upull(PullCtx pc, const unsigned char* cover, double* u4,
                                               int* stamp, int cycle, int fresh) {
  using T = typename Store<S>::T;
  const Geo& g = pc.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
  const bool want = cover[node] && (DENSE || fresh || stamp[node] != cycle);
  if (!DENSE && !want) return;
  T f[Lat<D>::Q];
  pull_node<D, S>(pc, i, j, k, f);
  const double4 out = upull_moments<D, S>(f, pc.base);
  if (!want) return;
  *reinterpret_cast<double4*>(u4 + 4 * (size_t)node) = out;
  stamp[node] = cycle;
}

// The dense gather of the reduced formats with the source rows of a 256-node
// chunk staged through LDS in 16-byte groups (stage_rows, as k_step_lds):
// the same values at every covered node, with 16-byte instead of 2- or
// 4-byte global loads.  Needs what k_step_lds needs (strides multiple of
// the group, state in pull form).
template <int D, typename S>
__global__ void __launch_bounds__(lds_chunk<S>()) k_upull_lds(PullCtx pc, const unsigned char* cover,
                                                              double* u4, int* stamp, int cycle) {
  constexpr int XC = lds_chunk<S>();
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  constexpr int G = LDS_GB / (int)sizeof(S);
  constexpr int W = XC + 2 * G;
  __shared__ __attribute__((aligned(16))) S lds[Q * W];
  S (*win)[W] = reinterpret_cast<S (*)[W]>(lds);
  const Geo g = pc.g;
  const int x0 = blockIdx.x * XC, j = blockIdx.y, k = blockIdx.z;
  const int xc = (g.nx - x0 < XC) ? g.nx - x0 : XC;
  stage_rows<D, S, XC>(g, pc.lay, static_cast<const S*>(pc.src), lds, x0, xc, j, k);
  __syncthreads();
  const int t = threadIdx.x;
  const int i = x0 + t;
  if (t >= xc) return;
  const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
  if (!cover[node]) return;
  T f[Q];
  pull_node_with<D, S>(pc, i, j, k, f, [&](int q, int, int, int) {
    return Store<S>::decode(win[q][t + G - L::ex(q)]);
  });
  *reinterpret_cast<double4*>(u4 + 4 * (size_t)node) = upull_moments<D, S>(f, pc.base);
  stamp[node] = cycle;
}

// A sub-step with moved markers on a kept marker ordering keeps the
// ordering's coverage and gathered velocities.  A marker whose base cell
// changed since the ordering was built may reach nodes the gather did not
// cover: one lane per marker finds the markers that left their ordered base
// cell, then the whole wavefront takes them one at a time, one lane per
// stencil node computing the velocity of a fluid node not yet gathered for
// this state (the same pull and moments as k_upull, stamped).  Two markers
// may compute the same node; they store the same values.
template <int D, typename S>
__global__ void __launch_bounds__(256) k_moved_fix(int n, const int* perm, const double* x, IbmParams ip,
                                                   PullCtx pc, const Base* bases, double* u4, int* stamp,
                                                   int cycle) {
  using T = typename Store<S>::T;
  const Geo& g = pc.g;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int i0 = 0, j0 = 0, k0 = 0;
  bool moved = false;
  if (p < n) {
    double lx, ly, lz;
    to_local(ip, D, x + 3 * (size_t)perm[p], lx, ly, lz);
    if (isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 && fabs(ly) < 1e9 && fabs(lz) < 1e9) {
      i0 = (int)floor(lx);
      j0 = (int)floor(ly);
      k0 = (D == 3) ? (int)floor(lz) : 0;
      const Base b = bases[p];
      moved = !(b.valid && b.i0 == i0 && b.j0 == j0 && b.k0 == k0);  // covered when ordered
    }
  }
  unsigned long long todo = __ballot(moved);
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1;
    const int mi = __shfl(i0, src), mj = __shfl(j0, src), mk = __shfl(k0, src);
    if (lane >= stencil_size<D>()) continue;
    const int a = mi - 1 + (lane & 3), bb = mj - 1 + ((lane >> 2) & 3);
    const int c = (D == 3) ? mk - 1 + (lane >> 4) : 0;
    if (a < 0 || a > g.nx + 1 || bb < 0 || bb > g.ny + 1 || (D == 3 && (c < 0 || c > g.nz + 1))) continue;
    int wd, wsd;
    if (node_type(pc.w, g, a, bb, c, wd, wsd) != 0) continue;
    const int i = wrap_ghost(a, g.nx), j = wrap_ghost(bb, g.ny), k = (D == 3) ? wrap_ghost(c, g.nz) : 0;
    const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
    if (stamp[node] == cycle) continue;
    T f[Lat<D>::Q];
    pull_node<D, S>(pc, i, j, k, f);
    *reinterpret_cast<double4*>(u4 + 4 * (size_t)node) = upull_moments<D, S>(f, pc.base);
    stamp[node] = cycle;
  }
}

// Interpolation sums of one marker over its stencil plane c = k0-1+vv (3D;
// the whole 4x4 stencil in 2D), IBM::interpolate's rules per node type
// (couplb_ibm.h:66-86): fluid nodes add vel(node)*d, bounce-back and
// free-slip nodes add d to the normalisation only, moving-wall nodes add
// the wall velocity; nodes beyond the ghost layer are skipped.  vel(node,
// slot, un) gives the velocity of fluid node `node`, stencil position
// slot = 4u + t of the plane.  s = {sum u,
// sum v, sum w, sum d over counted nodes, sum d over fluid nodes}.
// wall_vel = false drops the moving-wall velocity term (the weight still
// counts), which makes the map from node values to the marker value linear.
template <int D, typename Vel>
__device__ inline void interp_plane(const Geo& g, const Walls& w, bool fin, int i0, int j0, int k0,
                                    int vv, const double* wx, const double* wy, double wzv, Vel vel,
                                    double s[5], bool wall_vel = true) {
#pragma clang fp contract(off)
  s[0] = s[1] = s[2] = s[3] = s[4] = 0.0;
  const int c = (D == 3) ? k0 - 1 + vv : 0;
  if (!fin || (D == 3 && (c < 0 || c > g.nz + 1))) return;
  // two rows of the plane at a time: first the eight velocity loads (all in
  // flight together), then the sums in the reference's node order
#pragma unroll
  for (int u0 = 0; u0 < 4; u0 += 2) {
    int nt[2][4], wd[2][4], wsd[2][4];
    double un[2][4][3];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int b = j0 - 1 + u0 + h;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int a = i0 - 1 + t;
        const bool in = b >= 0 && b <= g.ny + 1 && a >= 0 && a <= g.nx + 1;
        nt[h][t] = in ? node_type(w, g, a, b, c, wd[h][t], wsd[h][t]) : -1;
        un[h][t][0] = un[h][t][1] = un[h][t][2] = 0.0;
        if (nt[h][t] == 0) {
          const unsigned node = (unsigned)wrap_ghost(a, g.nx) +
                                (unsigned)g.nx * ((unsigned)wrap_ghost(b, g.ny) +
                                                  (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0));
          vel(node, (u0 + h) * 4 + t, un[h][t]);
        }
      }
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int u = u0 + h;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const int tp = nt[h][t];
        if (tp < 0) continue;
        const double d = (D == 3) ? (wx[t] * wy[u]) * wzv : wx[t] * wy[u];
        if (tp == 0) {
          s[0] += un[h][t][0] * d;
          s[1] += un[h][t][1] * d;
          s[2] += un[h][t][2] * d;
          s[3] += d;
          s[4] += d;
        } else if (tp == 1 || tp == 3) {
          s[3] += d;
        } else if (tp == 2) {
          if (wall_vel) {
            const double* wvel = w.vel[wd[h][t]][wsd[h][t]];
            s[0] += wvel[0] * d;
            s[1] += wvel[1] * d;
            s[2] += ((D == 3) ? wvel[2] : 0.0) * d;
          }
          s[3] += d;
        }
      }
    }
  }
}

// combine the LPM lanes' plane sums in plane order 0, 1, 2, 3 (every lane
// of the group ends with the same totals)
template <int LPM>
__device__ inline void combine_planes(double s[5]) {
  if (LPM == 1) return;
  const int lane0 = (threadIdx.x & 63) & ~(LPM - 1);
  double t[5] = {0, 0, 0, 0, 0};
#pragma unroll
  for (int l = 0; l < LPM; l++)
#pragma unroll
    for (int r = 0; r < 5; r++) t[r] += __shfl(s[r], lane0 + l);
#pragma unroll
  for (int r = 0; r < 5; r++) s[r] = t[r];
}

// Markers in base-row order, four lanes per marker in 3D (one per z-plane
// of the 4x4x4 stencil).  IBM::interpolate (couplb_ibm.h:57-90) +
// penalty force (fix_couplb.cpp:779-785) + the data IBM::spread needs
// (couplb_ibm.h:92-133), stored as a MarkerRec.  The per-plane partial sums
// are combined in plane order, so results are deterministic; they differ
// from the reference's single running sum by rounding only.
// fmode 0: f_atom[idx] += F (device f); 1: f_atom[idx] = 0.0 + F (f taken as zero).
// Velocity cache (moving markers, later sub-steps of an LB interval): with
// ucache set, each lane keeps the velocities of its 16 stencil nodes (three
// doubles each, in stencil order) and cbase[p] the base cell and state they
// belong to; a marker whose base cell and state match reads them there
// instead of gathering them from u4 again.  The values depend only on the
// base cell and the state, so the cache stays valid when the marker
// ordering changes.
struct VelCache {
  double* u;   // value c of stencil slot q of lane l at u[(3q + c) * L + l]: one
               //   wavefront's 64 lanes read or write 64 consecutive doubles
  int4* base;  // per sorted marker: i0, j0, k0, state
  long L;      // lanes the cache holds (markers x lanes per marker)
  int cycle;   // current state
};

template <int D>
__global__ void __launch_bounds__(256) k_marker(int n, const int* perm, const double* x,
                                                const double* v, const double* m, double* f_atom,
                                                int fmode, IbmParams ip, Geo g, Walls w,
                                                const double* u4, MarkerRec* rec, double* ufc,
                                                double* fsrc, VelCache vc) {
#pragma clang fp contract(off)
  constexpr int LPM = (D == 3) ? 4 : 1;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = tid / LPM, vv = tid % LPM;
  const bool live = p < n;
  const int idx = live ? perm[p] : 0;
  double lx = 0, ly = 0, lz = 0;
  if (live) to_local(ip, D, x + 3 * (size_t)idx, lx, ly, lz);
  const bool fin = live && isfinite(lx) && isfinite(ly) && isfinite(lz) && fabs(lx) < 1e9 &&
                   fabs(ly) < 1e9 && fabs(lz) < 1e9;
  const int i0 = fin ? (int)floor(lx) : -100;
  const int j0 = fin ? (int)floor(ly) : -100;
  const int k0 = (fin && D == 3) ? (int)floor(lz) : 0;
  double wx[4], wy[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    wx[t] = delta1(ip.kern, lx - (double)(i0 - 1 + t));
    wy[t] = delta1(ip.kern, ly - (double)(j0 - 1 + t));
  }
  const double wzv = (D == 3) ? delta1(ip.kern, lz - (double)(k0 - 1 + vv)) : 1.0;
  bool cached = false;
  double* cl = nullptr;
  if (vc.u && fin) {
    const int4 cb = vc.base[p];
    cached = cb.x == i0 && cb.y == j0 && cb.z == k0 && cb.w == vc.cycle;
    cl = vc.u + (size_t)p * LPM + vv;
  }
  const size_t L = (size_t)vc.L;
  double s[5];
  interp_plane<D>(g, w, fin, i0, j0, k0, vv, wx, wy, wzv,
                  [&](unsigned node, int slot, double* un) {
                    if (cached) {
                      un[0] = cl[(3 * slot) * L];
                      un[1] = cl[(3 * slot + 1) * L];
                      un[2] = cl[(3 * slot + 2) * L];
                    } else {
                      const double4 q = *reinterpret_cast<const double4*>(u4 + 4 * (size_t)node);
                      un[0] = q.x; un[1] = q.y; un[2] = q.z;
                      if (cl) {
                        cl[(3 * slot) * L] = q.x;
                        cl[(3 * slot + 1) * L] = q.y;
                        cl[(3 * slot + 2) * L] = q.z;
                      }
                    }
                  }, s);
  if (cl && !cached && vv == 0) vc.base[p] = int4{i0, j0, k0, vc.cycle};
  combine_planes<LPM>(s);
  double ui = s[0], vi = s[1], wi = s[2], ws = s[3], wsf = s[4];
  double wz[4] = {1.0, 1.0, 1.0, 1.0};
  if (LPM > 1) {
    const int lane0 = (threadIdx.x & 63) & ~(LPM - 1);
#pragma unroll
    for (int l = 0; l < LPM; l++) wz[l] = __shfl(wzv, lane0 + l);
  }
  if (!live || vv != 0) return;
  double uf0 = 0, uf1 = 0, uf2 = 0;
  if (ws > DELTA_TOL) {
    const double iw = 1.0 / ws;
    uf0 = ui * iw; uf1 = vi * iw; uf2 = wi * iw;
  }
  ufc[3 * (size_t)p + 0] = uf0;
  ufc[3 * (size_t)p + 1] = uf1;
  ufc[3 * (size_t)p + 2] = uf2;
  const double* vp = v + 3 * (size_t)idx;
  const double mm = m[idx];
  const double Fx = mm * ip.xi_over_dtLB * (uf0 * ip.vel_scale - vp[0]);
  const double Fy = mm * ip.xi_over_dtLB * (uf1 * ip.vel_scale - vp[1]);
  const double Fz = mm * ip.xi_over_dtLB * (uf2 * ip.vel_scale - vp[2]);
  double* fa = f_atom + 3 * (size_t)idx;
  if (fmode == 0) { fa[0] += Fx; fa[1] += Fy; fa[2] += Fz; }
  else { fa[0] = 0.0 + Fx; fa[1] = 0.0 + Fy; fa[2] = 0.0 + Fz; }
  MarkerRec r;
#pragma unroll
  for (int t = 0; t < 4; t++) { r.wx[t] = wx[t]; r.wy[t] = wy[t]; r.wz[t] = wz[t]; }
  r.fx = -Fx * ip.inv_N / ip.force_scale;
  r.fy = -Fy * ip.inv_N / ip.force_scale;
  r.fz = -Fz * ip.inv_N / ip.force_scale;
  r.i0 = i0; r.j0 = j0; r.k0 = k0;
  r.valid = (fin && !(wsf < DELTA_TOL)) ? 1 : 0;
  r.sc = r.valid ? 1.0 / wsf : 0.0;
  rec[p] = r;
  if (fsrc) {  // compact copy of the reaction for the fixed-marker map
    fsrc[3 * (size_t)p + 0] = r.fx;
    fsrc[3 * (size_t)p + 1] = r.fy;
    fsrc[3 * (size_t)p + 2] = r.fz;
  }
}

// A later sub-step with unchanged positions in the same LB state: the
// stencil, the fluid velocity field and therefore u_f (ufc, written by
// k_marker) are the same as at the first sub-step, so only the penalty force
// (fix_couplb.cpp:779-785) is recomputed from the current v and m.  Its
// reaction is added into the first sub-step's record; spreading the summed
// reaction once equals spreading each sub-step's reaction (couplb_ibm.h:
// 125-130 is linear in the marker force) up to summation order.
//
// Sorted order (iperm null): thread p is sorted marker p, atom perm[p].
// Caller order (iperm: sorted position of each atom): thread t < n is atom
// lo + t, so that a range of atoms can run as soon as its arrays are on the
// device (the host-array sub-steps send them in chunks).
__global__ void __launch_bounds__(256) k_marker_force(int n, const int* perm, const double* v,
                                                      const double* m, double* f_atom, int fmode,
                                                      IbmParams ip, const double* ufc,
                                                      MarkerRec* rec, double* fsrc,
                                                      int lo = 0, const int* iperm = nullptr) {
#pragma clang fp contract(off)
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int idx = iperm ? lo + t : perm[t];
  const int p = iperm ? iperm[idx] : t;
  const double uf0 = ufc[3 * (size_t)p + 0], uf1 = ufc[3 * (size_t)p + 1], uf2 = ufc[3 * (size_t)p + 2];
  const double* vp = v + 3 * (size_t)idx;
  const double mm = m[idx];
  const double Fx = mm * ip.xi_over_dtLB * (uf0 * ip.vel_scale - vp[0]);
  const double Fy = mm * ip.xi_over_dtLB * (uf1 * ip.vel_scale - vp[1]);
  const double Fz = mm * ip.xi_over_dtLB * (uf2 * ip.vel_scale - vp[2]);
  double* fa = f_atom + 3 * (size_t)idx;
  if (fmode == 0) { fa[0] += Fx; fa[1] += Fy; fa[2] += Fz; }
  else { fa[0] = 0.0 + Fx; fa[1] = 0.0 + Fy; fa[2] = 0.0 + Fz; }
  MarkerRec& r = rec[p];
  r.fx = r.fx + -Fx * ip.inv_N / ip.force_scale;
  r.fy = r.fy + -Fy * ip.inv_N / ip.force_scale;
  r.fz = r.fz + -Fz * ip.inv_N / ip.force_scale;
  if (fsrc) {
    fsrc[3 * (size_t)p + 0] = r.fx;
    fsrc[3 * (size_t)p + 1] = r.fy;
    fsrc[3 * (size_t)p + 2] = r.fz;
  }
}

// iperm[perm[p]] = p
__global__ void __launch_bounds__(256) k_inv_perm(int n, const int* perm, int* iperm) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) iperm[perm[p]] = p;
}

}  // namespace couplb
