// kernels.hpp -- HIP kernels of the LB/IBM step for CDNA4 (gfx950).
//
// Device layout (interior nodes only, no ghost planes):
//   populations  f[q*pq + row*pr + i], row = j + ny*k, x fastest (Layout):
//                q planes (pq >= nx*ny*nz, pr = nx) or x-rows holding all
//                q together (pq >= nx, pr >= Q*pq); two buffers,
//                ping-ponged by the fused step.
//   The stored state is POST-COLLISION (f* of the last step) unless the
//   context is in "local" mode (right after create/import), when the buffer
//   holds the POST-STREAM populations the reference keeps between steps.
//   Periodic faces are index wraps; wall faces are coordinate tests; the
//   reference's ghost planes (couplb_lattice.h:82-104) are never stored.
#pragma once
#include <cstddef>

#include "lattice.hpp"

namespace couplb {

constexpr int MAX_SLOTS = 16;
// nodes of one 4-point IBM stencil (4^D)
template <int D> __host__ __device__ constexpr int stencil_size() { return D == 3 ? 64 : 16; }
constexpr double RHO_MIN = 1e-10;     // couplb_lattice.h:14
constexpr double DELTA_TOL = 1e-12;   // couplb_lattice.h:15

// Face configuration: type per face and the velocity of type-2 faces.
struct Walls {
  int type[3][2];
  double vel[3][2][3];
};

// One IBM marker contribution record, written by k_marker and read by the
// spreading (the row gather k_spread_rows, or the pair map built from it
// for fixed markers).  The weights are the 1-D delta values at the
// four stencil offsets -1..+2 of the base cell (i0, j0, k0) in the
// reference's ghost-inclusive coordinates (couplb_ibm.h:57-73).
struct MarkerRec {
  double wx[4], wy[4], wz[4];
  double sc;             // 1 / sum of fluid-node weights (couplb_ibm.h:123)
  double fx, fy, fz;     // -F*inv_N/force_scale (fix_couplb.cpp:787-792)
  int i0, j0, k0, valid;
};

static_assert(sizeof(MarkerRec) % 16 == 0 && offsetof(MarkerRec, fz) == 15 * sizeof(double),
              "the spreading loads a record's first 16 doubles as 16-byte groups");

struct Slot {
  const MarkerRec* rec;  // one record per marker, in the sorted marker order
};

// The record slots of an LB interval, as kernel arguments: records of slot
// s (one per marker, sorted marker order) are entries off[s] .. off[s + 1] - 1
// of their concatenation.  bases: when set, only the records whose base cell
// differs from the ordering's (the marker left it) are spread from here.
struct Base { int i0, j0, k0, valid; };  // a marker's base cell in the ordering
struct SlotSpan {
  int nslots;
  int off[MAX_SLOTS + 1];
  const MarkerRec* rec[MAX_SLOTS];
  const Base* bases;
};

struct Geo {
  int nx, ny, nz;        // interior extents
  int D;
  int nry;               // row-key stride in y: gy + 3 = ny + 5
  int nrows;             // number of row keys
};

// ---------------------------------------------------------------------------
// Ghost-inclusive node classification (couplb_boundary.h:12-35 with the
// x -> y -> z override order): returns the type of ghost-inclusive node
// (a, b, c); fills the wall dimension and side when it is a wall.
// ---------------------------------------------------------------------------
__device__ inline int node_type(const Walls& w, const Geo& g, int a, int b, int c,
                                int& wd, int& ws) {
  int t = 0;
  wd = -1; ws = 0;
  if (a == 0 && w.type[0][0] > 0) { t = w.type[0][0]; wd = 0; ws = 0; }
  if (a == g.nx + 1 && w.type[0][1] > 0) { t = w.type[0][1]; wd = 0; ws = 1; }
  if (b == 0 && w.type[1][0] > 0) { t = w.type[1][0]; wd = 1; ws = 0; }
  if (b == g.ny + 1 && w.type[1][1] > 0) { t = w.type[1][1]; wd = 1; ws = 1; }
  if (g.D == 3) {
    if (c == 0 && w.type[2][0] > 0) { t = w.type[2][0]; wd = 2; ws = 0; }
    if (c == g.nz + 1 && w.type[2][1] > 0) { t = w.type[2][1]; wd = 2; ws = 1; }
  }
  return t;
}

// ghost-inclusive index (1..n interior, 0 / n+1 ghosts) -> interior 0-based
// index of the periodic image
__device__ inline int wrap_ghost(int a, int n) {
  return a == 0 ? n - 1 : (a == n + 1 ? 0 : a - 1);
}

// ---------------------------------------------------------------------------
// Pull of the 19 (9) populations arriving at interior node (i, j, k).
// Replaces Streaming::stream (couplb_streaming.h:93-132) applied to the
// post-collision populations src.  Rules per source node type:
//   0 (fluid or periodic image) f[q][s]; 1 no-slip f[opp q][n];
//   2 moving   f[opp q][n] + 2 w rho_w (e_q . u_wall)/cs^2,
//              rho_w = max(rho_coll[n], rho_floor);
//   3 free-slip f[reflect_d q][n]; 4 open f[q][n].
// ---------------------------------------------------------------------------
// Element (q, x, row) of a population buffer: q*pq + row*pr + x.
struct Layout {
  size_t pq, pr;
  __host__ __device__ size_t at(int q, unsigned row, unsigned x) const {
    return (size_t)q * pq + (size_t)row * pr + x;
  }
};

struct PullCtx {
  const void* src;
  Layout lay;
  Geo g;
  Walls w;
  const double* rho_coll;  // density of the collision that made src
  double rho_floor;
  int local;               // src already holds post-stream populations
  double base;             // stored offset per unit weight: rho0 for the
                           // deviation formats, 0 for FP64
  double rho0;
  float scale, inv_scale;  // storage scale of reduced formats
};

// fluid(q, si, sj, sk) returns the stored value of plane q at fluid source
// node (si, sj, sk); the other rules read the node itself from global memory
// lay, src, rho_coll: the state the wall rules read and the density of the
// collision that made it (pc.lay, pc.src, pc.rho_coll in global memory for
// the single-step kernels; LDS copies in k_steps_fused)
// SCALED: leave the reduced formats' values in their stored scale (the LB
// step works in it); otherwise divide by it
template <int D, typename S, bool SCALED = false, typename Fluid>
__device__ inline void pull_node_with(const PullCtx& pc, const Layout& lay, const S* src,
                                      const double* rho_coll, int i, int j, int k,
                                      typename Store<S>::T (&f)[Lat<D>::Q], Fluid fluid) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const int nx = pc.g.nx, ny = pc.g.ny, nz = pc.g.nz;
  const unsigned row = (unsigned)j + (unsigned)ny * (unsigned)k;
  const unsigned node = (unsigned)i + (unsigned)nx * row;
  if (pc.local) {
#pragma unroll
    for (int q = 0; q < Q; q++) {
      f[q] = Store<S>::load(src + lay.at(q, row, i));
      if (Store<S>::deviation && !SCALED) f[q] *= (T)pc.inv_scale;
    }
    return;
  }
  // block-uniform y/z neighbour rows (wrapped) and wall flags
  const int jm = (j == 0) ? ny - 1 : j - 1;       // source row for ey = +1
  const int jp = (j == ny - 1) ? 0 : j + 1;       // source row for ey = -1
  const bool ylo = (j == 0) && pc.w.type[1][0] > 0;
  const bool yhi = (j == ny - 1) && pc.w.type[1][1] > 0;
  int km = 0, kp = 0;
  bool zlo = false, zhi = false;
  if (D == 3) {
    km = (k == 0) ? nz - 1 : k - 1;
    kp = (k == nz - 1) ? 0 : k + 1;
    zlo = (k == 0) && pc.w.type[2][0] > 0;
    zhi = (k == nz - 1) && pc.w.type[2][1] > 0;
  }
  const int im = (i == 0) ? nx - 1 : i - 1;
  const int ip = (i == nx - 1) ? 0 : i + 1;
  const bool xlo = (i == 0) && pc.w.type[0][0] > 0;
  const bool xhi = (i == nx - 1) && pc.w.type[0][1] > 0;
  const bool any_moving = (xlo && pc.w.type[0][0] == 2) || (xhi && pc.w.type[0][1] == 2) ||
                          (ylo && pc.w.type[1][0] == 2) || (yhi && pc.w.type[1][1] == 2) ||
                          (zlo && pc.w.type[2][0] == 2) || (zhi && pc.w.type[2][1] == 2);
  if (!(xlo || xhi || ylo || yhi || zlo || zhi)) {
    // no wall face next to this node (most of the grid): every population
    // comes from its fluid or periodic source, as in the loop below
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int ex = L::ex(q), ey = L::ey(q), ez = L::ez(q);
      const int si = ex == 1 ? im : (ex == -1 ? ip : i);
      const int sj = ey == 1 ? jm : (ey == -1 ? jp : j);
      const int sk = (D == 3) ? (ez == 1 ? km : (ez == -1 ? kp : k)) : 0;
      f[q] = fluid(q, si, sj, sk);
      if (Store<S>::deviation && !SCALED) f[q] *= (T)pc.inv_scale;
    }
    return;
  }
  T rw = 0;
  if (any_moving) {
    const double r = rho_coll[node];
    rw = (T)((r < pc.rho_floor) ? pc.rho_floor : r);
  }
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int ex = L::ex(q), ey = L::ey(q), ez = L::ez(q);
    // wall lookup with z > y > x priority (later set_walls_* override)
    int wt = 0, wd = 0, wsd = 0;
    if (ex == 1 && xlo) { wt = pc.w.type[0][0]; wd = 0; wsd = 0; }
    if (ex == -1 && xhi) { wt = pc.w.type[0][1]; wd = 0; wsd = 1; }
    if (ey == 1 && ylo) { wt = pc.w.type[1][0]; wd = 1; wsd = 0; }
    if (ey == -1 && yhi) { wt = pc.w.type[1][1]; wd = 1; wsd = 1; }
    if (D == 3) {
      if (ez == 1 && zlo) { wt = pc.w.type[2][0]; wd = 2; wsd = 0; }
      if (ez == -1 && zhi) { wt = pc.w.type[2][1]; wd = 2; wsd = 1; }
    }
    const int si = ex == 1 ? im : (ex == -1 ? ip : i);
    const int sj = ey == 1 ? jm : (ey == -1 ? jp : j);
    const int sk = (D == 3) ? (ez == 1 ? km : (ez == -1 ? kp : k)) : 0;
    if (wt == 0) {
      f[q] = fluid(q, si, sj, sk);
    } else {
      int qs = q;
      if (wt == 1 || wt == 2) qs = L::opp(q);
      else if (wt == 3) qs = (wd == 0) ? L::reflect(0, q) : (wd == 1 ? L::reflect(1, q) : L::reflect(2, q));
      f[q] = Store<S>::load(src + lay.at(qs, row, i));
      if (wt == 2) {
        const double* uw = pc.w.vel[wd][wsd];
        if (Store<S>::deviation) {
          // 2 w rho_w eu / cs^2 with cs^2 = 1/3; stored values carry `scale`
          const T eu = (T)(ex * uw[0] + ey * uw[1] + ez * uw[2]);
          f[q] += (T)(6.0 * L::w(L::opp(q))) * rw * eu * (T)pc.scale;
        } else {
#pragma clang fp contract(off)
          // f[opp] + 2.0*w[opp]*rw*eu/cs2 in the reference's order
          double eu = 0.0;
          if (ex) eu = ex * uw[0];
          if (ey) eu = (ex ? eu + ey * uw[1] : ey * uw[1]);
          if (ez) eu = ((ex || ey) ? eu + ez * uw[2] : ez * uw[2]);
          f[q] = f[q] + div_cs2(2.0 * L::w(L::opp(q)) * (double)rw * eu);
        }
      }
    }
    if (Store<S>::deviation && !SCALED) f[q] *= (T)pc.inv_scale;
  }
}

template <int D, typename S, bool SCALED = false, typename Fluid>
__device__ inline void pull_node_with(const PullCtx& pc, int i, int j, int k,
                                      typename Store<S>::T (&f)[Lat<D>::Q], Fluid fluid) {
  pull_node_with<D, S, SCALED>(pc, pc.lay, static_cast<const S*>(pc.src), pc.rho_coll, i, j, k, f, fluid);
}

// TX: the x-shifted pulls (e_x != 0) as ordinary loads instead of
// nontemporal ones, so that the 128-byte line a wavefront shares with its
// x-neighbour stays in L2 for it.  Measured faster only in the FP64 step
// with the fixed-marker map (about 1 %); slower in the plain step and the
// velocity gather.
template <int D, typename S, bool SCALED = false, bool TX = false>
__device__ inline void pull_node(const PullCtx& pc, int i, int j, int k,
                                 typename Store<S>::T (&f)[Lat<D>::Q]) {
  const S* src = static_cast<const S*>(pc.src);
  const int ny = pc.g.ny;
  pull_node_with<D, S, SCALED>(pc, i, j, k, f, [&](int q, int si, int sj, int sk) {
    const S* p = src + pc.lay.at(q, (unsigned)sj + (unsigned)ny * (unsigned)sk, (unsigned)si);
    if (TX && Lat<D>::ex(q) != 0) return Store<S>::decode(*p);
    return Store<S>::load(p);
  });
}

// ---------------------------------------------------------------------------
// The fused step: pull (stream of the previous step) -> IBM force (the
// node's pair-run sum for fixed markers, else the spread node force) ->
// gravity -> macroscopic with Guo half-force -> BGK + Guo collision -> store.
// Replaces do_lbm_step (fix_couplb.cpp:659-669): apply_external_force
// (680-708), compute_macroscopic(true) (couplb_lattice.h:155-202),
// BGK::collide (couplb_collision.h:18-62), exchange + stream
// (couplb_streaming.h:82-139) and clear_forces (couplb_lattice.h:243-247).
// ---------------------------------------------------------------------------
// a + b*c; the FP64 path rounds the product separately, as the reference
template <typename T>
__device__ inline T add_prod(T a, T b, T c) { return a + b * c; }
template <>
__device__ inline double add_prod<double>(double a, double b, double c) {
#pragma clang fp contract(off)
  return a + b * c;
}

// Signed sum of the nonzero terms of c . v for a lattice direction: equal in
// value to the reference's cx*v0 + cy*v1 + cz*v2 (the omitted terms are
// products with 0 and do not change a sum).
template <int D>
__device__ inline double edot(int q, double v0, double v1, double v2) {
#pragma clang fp contract(off)
  using L = Lat<D>;
  const int ex = L::ex(q), ey = L::ey(q), ez = L::ez(q);
  double s = 0.0;
  bool any = false;
  if (ex) { s = ex > 0 ? v0 : -v0; any = true; }
  if (ey) { s = any ? (ey > 0 ? s + v1 : s - v1) : (ey > 0 ? v1 : -v1); any = true; }
  if (ez) { s = any ? (ez > 0 ? s + v2 : s - v2) : (ez > 0 ? v2 : -v2); any = true; }
  return s;
}

// First moments sum_q e_q f_q with every product, in q order, as
// compute_macroscopic sums them (couplb_lattice.h:167-173): used where a
// population is not finite, so that 0 * Inf = NaN reaches the same
// components as in the reference.
template <int D, typename V>
__device__ inline void moments_full(const V* f, double& jx, double& jy, double& jz) {
#pragma clang fp contract(off)
  using L = Lat<D>;
  jx = jy = jz = 0.0;
#pragma unroll
  for (int q = 0; q < L::Q; q++) {
    const double fv = (double)f[q];
    jx += (double)L::ex(q) * fv;
    jy += (double)L::ey(q) * fv;
    jz += (double)L::ez(q) * fv;
  }
}

// e_q . v for the reduced-precision collision: the nonzero terms only, so no
// multiplications by 0 or +-1 are issued (they cannot be dropped by the
// compiler under IEEE rules)
template <int D, typename T>
__device__ inline T esum(int q, T v0, T v1, T v2) {
  using L = Lat<D>;
  const int ex = L::ex(q), ey = L::ey(q), ez = L::ez(q);
  T s = (T)0;
  bool any = false;
  if (ex) { s = ex > 0 ? v0 : -v0; any = true; }
  if (ey) { s = any ? (ey > 0 ? s + v1 : s - v1) : (ey > 0 ? v1 : -v1); any = true; }
  if (D == 3 && ez) { s = any ? (ez > 0 ? s + v2 : s - v2) : (ez > 0 ? v2 : -v2); }
  return s;
}

// BGK + Guo in FP64 with the reference's expression order and rounding
// (couplb_lattice.h:196-200 velocity, couplb_collision.h:36-58 collision):
// the result equals the reference's to the last bit.  With a non-finite
// velocity or force every output of the node is non-finite here as in the
// reference (u.u or u.F is), but may be an infinity where the reference's
// is NaN: the dot products skip the zero components whose products turn an
// infinity into NaN there.
template <int D>
__device__ inline void collide_exact(const double (&f)[Lat<D>::Q], double r, double su, double sv,
                                     double sw, double F0, double F1, double F2, double om,
                                     double* dst, size_t pq, size_t at0) {
#pragma clang fp contract(off)
  using L = Lat<D>;
  constexpr int Q = L::Q;
  const double inv_r = 1.0 / r;
  const double u0 = (su + 0.5 * F0) * inv_r;
  const double u1 = (sv + 0.5 * F1) * inv_r;
  const double u2 = (sw + 0.5 * F2) * inv_r;
  const double half_om = 1.0 - 0.5 * om;
  const double uu = u0 * u0 + u1 * u1 + u2 * u2;
  const double uf = u0 * F0 + u1 * F1 + u2 * F2;
  const double uu_term = div_cs2(0.5 * uu);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double w = L::w(q);
    const double eu = edot<D>(q, u0, u1, u2);
    const double ef = edot<D>(q, F0, F1, F2);
    const double feq = w * r * (1.0 + div_cs2(eu) + div_cs4(0.5 * eu * eu) - uu_term);
    const double Si = w * half_om * (div_cs2(ef - uf) + div_cs4(eu * ef));
    const double fn = f[q] + (-om * (f[q] - feq) + Si);
    Store<double>::store(dst + q * pq + at0, fn);
  }
}

struct StepArgs {
  PullCtx pc;
  void* dst;
  double omega;
  double g[3];
  int grav_mode;        // 0 none, 1 current density, 2 density array (stale)
  int write_rho;        // 1: store the collision density at every node
  int has_moving;       // store the density at nodes next to type-2 faces
  double* rho_out;      // may alias pc.rho_coll (same node read then written)
  int has_ibm;          // spread force of the last cycle present
  const unsigned char* flag;
  const double* F4;    // node force (fx, fy, fz, unused) where flag is set
  // fixed-marker map: instead of F4, the step sums the node's run of
  // (marker, weight) pairs itself: the runs of a 64-node segment of an x-row
  // start at pseg[segment] and follow each other in node order, plen giving
  // their lengths (pair_run); pp/pw the sorted pairs, fsrc three reaction
  // values per sorted marker
  int pair_mode;        // 0 F4 at flagged nodes, 1 fixed-marker map, 2 moving-marker map
  const unsigned short* plen;  // pair run length per node (0: no pairs)
  const int* pseg;             // first pair of each 64-node segment of an x-row
  const int* pp;
  const double* pw;
  const double* fsrc;
  // moving-marker map (pair_mode 2): pcon holds the force of every (marker,
  // stencil slot) summed over the sub-steps (four doubles each, the last
  // unused), pidx the sorted pairs' (marker, stencil slot) entries; records
  // that left the map's base cells were spread into F4 at the nodes where
  // flag2 holds the step's tag
  const double* pcon;
  const int* pidx;
  const unsigned char* flag2;
  int has_f2;           // pair_mode 2: F4 applies where flag2 holds this tag (1..255)
};

// pairs in flight per round of the run sums in the reduced formats' steps
#ifndef COUPLB_RUN_U_REDUCED
#define COUPLB_RUN_U_REDUCED 2
#endif
constexpr int RUN_U_REDUCED = COUPLB_RUN_U_REDUCED;

// Sum of one node's run of the moving-marker map, sum_f con[idx[f]] in pair
// order (con: four doubles per entry, the last unused).
template <int U = 4>
__device__ inline void run_sum_con(int s, int len, const int* idx, const double* con, double& ax, double& ay,
                                   double& az) {
#pragma clang fp contract(off)
  ax = ay = az = 0.0;
  const double4* c4 = reinterpret_cast<const double4*>(con);
  int f = s;
  const int end = s + len;
  // U entries in flight per round; the sums are taken in pair order either way
  for (; U > 1 && f + U <= end; f += U) {
    int q[U];
    double4 v[U];
#pragma unroll
    for (int i = 0; i < U; i++) q[i] = idx[f + i];
#pragma unroll
    for (int i = 0; i < U; i++) v[i] = c4[q[i]];
#pragma unroll
    for (int i = 0; i < U; i++) {
      ax += v[i].x;
      ay += v[i].y;
      az += v[i].z;
    }
  }
  for (; f < end; f++) {
    const double4 v = c4[idx[f]];
    ax += v.x;
    ay += v.y;
    az += v.z;
  }
}

// Sum of one node's run of the fixed-marker map, sum_f src[pp[f]] * pw[f]
// in pair order (couplb_ibm.h:125-130: F_node += F_marker * w); src holds
// three values per sorted marker.  U pairs are loaded per round; the sums
// are the same.
template <int U = 4>
__device__ inline void run_sum(int s, int len, const int* pp, const double* pw, const double* src,
                               double& ax, double& ay, double& az) {
#pragma clang fp contract(off)
  ax = ay = az = 0.0;
  int f = s;
  const int end = s + len;
  for (; U > 1 && f + U <= end; f += U) {
    int q[U];
    double w[U], v[U][3];
#pragma unroll
    for (int i = 0; i < U; i++) { q[i] = pp[f + i]; w[i] = pw[f + i]; }
#pragma unroll
    for (int i = 0; i < U; i++)
#pragma unroll
      for (int c = 0; c < 3; c++) v[i][c] = src[3 * (size_t)q[i] + c];
#pragma unroll
    for (int i = 0; i < U; i++) {
      ax += v[i][0] * w[i];
      ay += v[i][1] * w[i];
      az += v[i][2] * w[i];
    }
  }
  for (; f < end; f++) {
    const int q = pp[f];
    const double w = pw[f];
    ax += src[3 * (size_t)q + 0] * w;
    ay += src[3 * (size_t)q + 1] * w;
    az += src[3 * (size_t)q + 2] * w;
  }
}

// Everything of the step after the pull: IBM force, gravity, macroscopic
// with the Guo half-force, density bookkeeping, collision.  out(q, v) stores
// the post-collision value of plane q (already scaled for the reduced
// formats).
// lay, dst: the buffer collide_exact (FP64) writes (the reduced formats
// store through out); rho_in, rho_out: the density array read for gravity
// and the one written (a.pc.lay, a.dst, a.pc.rho_coll, a.rho_out in the
// single-step kernels).  PM 1: flagged nodes take their force from their
// run of the fixed-marker map (pair_run, a.pp, a.pw, a.fsrc) instead of a.F4;
// PM 2: from their run of the moving-marker map (a.pcon), plus a.F4 where
// a.flag2 is set; separate instantiations, so the other steps keep their
// registers.
// The pair run (first pair, length) of interior node (i, row) of a pair map:
// its segment's first pair plus the lengths of the runs of the lanes before
// it.  Wavefront-collective: every lane of the wavefront calls it, the 64
// lanes hold 64 consecutive x of one row starting at a multiple of 64, and
// lanes past the row's end pass live = false.
__device__ inline int2 pair_run(const StepArgs& a, int i, unsigned row, bool live) {
  const Geo& g = a.pc.g;
  const unsigned nsx = ((unsigned)g.nx + 63u) >> 6;
  const int len = live ? (int)a.plen[(size_t)i + (size_t)g.nx * row] : 0;
  int incl = len;
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  const int i0 = i & ~63;  // the same segment for every lane
  const int first = (i0 < g.nx) ? a.pseg[(size_t)row * nsx + (unsigned)(i0 >> 6)] : 0;
  return int2{first + incl - len, len};
}

// IBM force spread by the last cycle's markers, present only at flagged
// nodes: the node's pair run (PM 1: fixed-marker map; PM 2: moving-marker
// map, plus the tile-spread F4 where flag2 holds the step's tag) or the tile
// spread's F4 (PM 0).
template <int D, typename S, int PM>
__device__ inline void ibm_node_force(const StepArgs& a, unsigned node, int2 r, typename Store<S>::T& Fx,
                                      typename Store<S>::T& Fy, typename Store<S>::T& Fz) {
  using T = typename Store<S>::T;
  Fx = 0;
  Fy = 0;
  Fz = 0;
  if constexpr (PM == 2) {
    if (a.has_ibm) {
      double ax = 0.0, ay = 0.0, az = 0.0;
      if (r.y > 0) run_sum_con<Store<S>::deviation ? RUN_U_REDUCED : 4>(r.x, r.y, a.pidx, a.pcon, ax, ay, az);
      if (a.has_f2 && a.flag2[node] == (unsigned char)a.has_f2) {
#pragma clang fp contract(off)
        const double4 F = *reinterpret_cast<const double4*>(a.F4 + 4 * (size_t)node);
        ax += F.x;
        ay += F.y;
        az += F.z;
      }
      Fx = (T)ax;
      Fy = (T)ay;
      Fz = (T)az;
    }
  } else if constexpr (PM == 1) {
    if (a.has_ibm && r.y > 0) {
      double ax, ay, az;
      // FP64: four pairs in flight (measured faster); the reduced formats
      // keep their occupancy with fewer
      run_sum<Store<S>::deviation ? RUN_U_REDUCED : 4>(r.x, r.y, a.pp, a.pw, a.fsrc, ax, ay, az);
      Fx = (T)ax;
      Fy = (T)ay;
      Fz = (T)az;
    }
  } else if (a.has_ibm && a.flag[node]) {
    const double4 F = *reinterpret_cast<const double4*>(a.F4 + 4 * (size_t)node);
    Fx = (T)F.x;
    Fy = (T)F.y;
    Fz = (T)F.z;
  }
}

template <int D, typename S, int PM = 0, typename Out>
__device__ inline void step_node(const StepArgs& a, const Layout& lay, S* dst, const double* rho_in,
                                 double* rho_out, int i, int j, int k, unsigned node,
                                 typename Store<S>::T (&f)[Lat<D>::Q], Out out,
                                 const typename Store<S>::T* Fpre = nullptr, int2 run = int2{0, 0}) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const Geo& g = a.pc.g;
  // IBM force (ibm_node_force), or the one the caller loaded beforehand
  T Fx, Fy, Fz;
  if (Fpre) {
    Fx = Fpre[0];
    Fy = Fpre[1];
    Fz = Fpre[2];
  } else {
    ibm_node_force<D, S, PM>(a, node, run, Fx, Fy, Fz);
  }
  // density (deviation form for the reduced formats: A = rho - base).  The
  // reduced formats' populations arrive in their stored scale s (a power of
  // two), so the sums are s times the unscaled ones, exactly
  T A = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    A += f[q];
    if (L::ex(q) == 1) jx += f[q]; else if (L::ex(q) == -1) jx -= f[q];
    if (L::ey(q) == 1) jy += f[q]; else if (L::ey(q) == -1) jy -= f[q];
    if (D == 3) { if (L::ez(q) == 1) jz += f[q]; else if (L::ez(q) == -1) jz -= f[q]; }
  }
  if constexpr (!Store<S>::deviation) {
    // a non-finite population: the moments with every product, as the
    // reference forms them (the sums above skip the zero components)
    if (!isfinite(A)) moments_full<D>(f, jx, jy, jz);
  }
  T As = A;  // A in the stored scale
  if constexpr (Store<S>::deviation) {
    const T is = (T)a.pc.inv_scale;
    A *= is;
    jx *= is;
    jy *= is;
    jz *= is;
  }
  const T base = (T)a.pc.base;
  const T mass_in = As;  // the node's stored populations summed (stored scale)
  T rho = A + base;
  // couplb_lattice.h:178-186: rho <= RHO_MIN, then any non-finite value
  // (NaN, +Inf) is clamped to RHO_MIN
  bool clamped = false;
  if (rho <= (T)RHO_MIN || !isfinite(rho)) {
    clamped = true;
    rho = (T)RHO_MIN;
    A = rho - base;
    As = A * (T)a.pc.scale;
  }
  // gravity: F += g * rho_array (fix_couplb.cpp:706)
  if (a.grav_mode) {
    const T rg = (a.grav_mode == 1) ? rho : (T)rho_in[node];
    Fx = add_prod(Fx, (T)a.g[0], rg);
    Fy = add_prod(Fy, (T)a.g[1], rg);
    Fz = add_prod(Fz, (T)a.g[2], rg);
  }
  if (a.write_rho) rho_out[node] = (double)rho;
  else if (a.has_moving) {
    const bool wall_adj =
        (i == 0 && a.pc.w.type[0][0] == 2) || (i == g.nx - 1 && a.pc.w.type[0][1] == 2) ||
        (j == 0 && a.pc.w.type[1][0] == 2) || (j == g.ny - 1 && a.pc.w.type[1][1] == 2) ||
        (D == 3 && ((k == 0 && a.pc.w.type[2][0] == 2) || (k == g.nz - 1 && a.pc.w.type[2][1] == 2)));
    if (wall_adj) rho_out[node] = (double)rho;
  }
  if constexpr (!Store<S>::deviation) {
    collide_exact<D>(f, rho, jx, jy, jz, Fx, Fy, Fz, a.omega, dst, lay.pq,
                     lay.at(0, (unsigned)j + (unsigned)g.ny * (unsigned)k, (unsigned)i));
    (void)out;
  } else {
    const T inv_r = (T)1 / rho;
    const T ux = (jx + (T)0.5 * Fx) * inv_r;
    const T uy = (jy + (T)0.5 * Fy) * inv_r;
    const T uz = (jz + (T)0.5 * Fz) * inv_r;
    const T om = (T)a.omega;
    const T hom = (T)1 - (T)0.5 * om;
    const T uu = ux * ux + uy * uy + uz * uz;
    const T uf = ux * Fx + uy * Fy + uz * Fz;
    const T scale = (T)a.pc.scale;
    // The populations stay in the stored scale s: every term below is s
    // times the unscaled one (s a power of two, so each product and sum is
    // the unscaled result times s exactly).
    // BGK + Guo per direction (couplb_collision.h:36-58), in deviation form:
    //   feq - w*base = w*(A + rho*(3eu + 4.5eu^2 - 1.5uu))   (couplb_lattice.h:137-144)
    //   S = w*(1-om/2)*(3(e.F - u.F) + 9(e.u)(e.F))
    //   f' = f - om*(f - feq) + S
    // The parts even in e (sym = A + rho*(4.5eu^2 - 1.5uu), Ssym =
    // 9 hom eu ef - 3 hom uf) are shared by q and opp(q); the odd parts
    // (asym = 3 rho eu, Sasym = 3 hom ef) change sign: one evaluation per
    // pair of opposite directions.  f - feq stays one difference, as in the
    // reference's form, so the non-equilibrium part keeps its precision.
    //
    // Mass: BGK with Guo forcing keeps sum_q f exactly in exact arithmetic.
    // The rest population (q = 0, which does not stream) is set to the
    // node's incoming mass minus the other populations as they are stored
    // after rounding, so the stored node mass changes only by the rounding
    // of that one value: not at all for the 16-bit fixed point, whose
    // stored values are integers in this scale (exact float sums).  A
    // clamped density changes the mass as in the reference: then the rest
    // population takes the collision formula.
    const T uu15 = (T)1.5 * uu, rho_s = rho * scale, rho3 = ((T)3 * rho) * scale;
    const T hom3 = ((T)3 * hom) * scale, hom9 = ((T)9 * hom) * scale, huf3 = hom3 * uf;
    T moved = 0, rest = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int o = L::opp(q);
      if (o < q) continue;
      const T eu = esum<D, T>(q, ux, uy, uz);
      const T ef = esum<D, T>(q, Fx, Fy, Fz);
      const T wq = (T)L::w(q);
      const T sym = As + rho_s * ((T)4.5 * eu * eu - uu15);
      const T Ssym = hom9 * eu * ef - huf3;
      if (o == q) {
        rest = f[q] - om * (f[q] - wq * sym) + wq * Ssym;
        continue;
      }
      const T asym = rho3 * eu, Sasym = hom3 * ef;
      const T vq = Store<S>::round(f[q] - om * (f[q] - wq * (sym + asym)) + wq * (Ssym + Sasym));
      const T vo = Store<S>::round(f[o] - om * (f[o] - wq * (sym - asym)) + wq * (Ssym - Sasym));
      out(q, vq);
      out(o, vo);
      moved += vq;
      moved += vo;
    }
    out(0, clamped ? rest : mass_in - moved);
  }
}

template <int D, typename S, int PM = 0, typename Out>
__device__ inline void step_node(const StepArgs& a, int i, int j, int k, unsigned node,
                                 typename Store<S>::T (&f)[Lat<D>::Q], Out out,
                                 const typename Store<S>::T* Fpre = nullptr, int2 run = int2{0, 0}) {
  step_node<D, S, PM>(a, a.pc.lay, static_cast<S*>(a.dst), a.pc.rho_coll, a.rho_out, i, j, k, node, f, out, Fpre,
                      run);
}

template <int D, typename S, int PM = 0>
__global__ void __launch_bounds__(256) k_step(StepArgs a) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const Geo& g = a.pc.g;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  const unsigned row = (unsigned)j + (unsigned)g.ny * (unsigned)k;
  // the node's pair run, computed by the whole wavefront before its lanes
  // past the row's end leave
  int2 run = int2{0, 0};
  if (PM != 0 && a.has_ibm) run = pair_run(a, i, row, i < g.nx);
  if (i >= g.nx) return;
  const unsigned node = (unsigned)i + (unsigned)g.nx * row;
  T f[Q];
  pull_node<D, S, true, PM == 1>(a.pc, i, j, k, f);
  S* dst = static_cast<S*>(a.dst);
  const size_t at0 = a.pc.lay.at(0, row, (unsigned)i);
  step_node<D, S, PM>(a, i, j, k, node, f,
                        [&](int q, T v) { Store<S>::store(dst + q * a.pc.lay.pq + at0, v); }, nullptr, run);
}

// nsteps LB steps of a small grid in one launch of one workgroup, with the
// state and the density array held in LDS: a step of a grid this small
// takes less time than a kernel launch, so a loop over launches would be
// bound by launch latency.  Two LDS state buffers in the row layout without
// padding (pq = nx, pr = Q*nx) and the density array, filled from a.pc.src
// and a.pc.rho_coll, then one barrier per step, and the last state and the
// density written back to the buffer the same steps launched one by one
// would have ended in (a.pc.src for an even count, a.dst for an odd one)
// and to a.rho_out (= a.pc.rho_coll).  The steps are those after the first
// of a run (state in pull form; gravity, if any, from the density the step
// before wrote), the same for every step but the buffers; the per-node
// arithmetic is k_step's, so the bits are k_step's.
constexpr int FUSED_THREADS = 1024;
template <int D, typename S>
__host__ __device__ constexpr size_t fused_state_elems(size_t nodes) { return nodes * (size_t)Lat<D>::Q; }
template <int D, typename S>
__host__ __device__ constexpr size_t fused_lds_bytes(size_t nodes) {
  return (2 * fused_state_elems<D, S>(nodes) * sizeof(S) + 7) / 8 * 8 + nodes * sizeof(double);
}

template <int D, typename S>
__global__ void __launch_bounds__(FUSED_THREADS) k_steps_fused(StepArgs a, int nsteps) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  extern __shared__ __attribute__((aligned(16))) unsigned char fused_lds[];
  const unsigned nx = (unsigned)a.pc.g.nx, ny = (unsigned)a.pc.g.ny;
  const unsigned N = nx * ny * (unsigned)(D == 3 ? a.pc.g.nz : 1);
  const Layout ll{nx, (size_t)Q * nx};
  const size_t ne = fused_state_elems<D, S>(N);
  S* A = reinterpret_cast<S*>(fused_lds);
  S* B = A + ne;
  double* R = reinterpret_cast<double*>(fused_lds + (2 * ne * sizeof(S) + 7) / 8 * 8);
  const S* gsrc = static_cast<const S*>(a.pc.src);
  for (unsigned node = threadIdx.x; node < N; node += blockDim.x) {
    const unsigned row = node / nx, i = node - row * nx;
#pragma unroll
    for (int q = 0; q < Q; q++) A[ll.at(q, row, i)] = gsrc[a.pc.lay.at(q, row, i)];
    R[node] = a.pc.rho_coll[node];
  }
  for (int s = 0; s < nsteps; s++) {
    __syncthreads();
    const S* src = (s & 1) ? B : A;
    S* dst = (s & 1) ? A : B;
    for (unsigned node = threadIdx.x; node < N; node += blockDim.x) {
      const unsigned row = node / nx, i = node - row * nx;
      const unsigned j = row % ny, k = row / ny;
      T f[Q];
      pull_node_with<D, S, true>(a.pc, ll, src, R, (int)i, (int)j, (int)k, f, [&](int q, int si, int sj, int sk) {
        return Store<S>::load(src + ll.at(q, (unsigned)sj + ny * (unsigned)sk, (unsigned)si));
      });
      const size_t at0 = ll.at(0, row, i);
      step_node<D, S>(a, ll, dst, R, R, (int)i, (int)j, (int)k, node, f,
                      [&](int q, T v) { Store<S>::store(dst + q * ll.pq + at0, v); });
    }
  }
  __syncthreads();
  // the last state, and the one before it in the other buffer, where the
  // same steps launched one by one would leave them (the state before the
  // last is what a checkpoint's ghost planes are made of)
  const S* fin = (nsteps & 1) ? B : A;
  const S* before = (nsteps & 1) ? A : B;
  S* gdst = static_cast<S*>((nsteps & 1) ? a.dst : const_cast<void*>(a.pc.src));
  S* gother = static_cast<S*>((nsteps & 1) ? const_cast<void*>(a.pc.src) : a.dst);
  for (unsigned node = threadIdx.x; node < N; node += blockDim.x) {
    const unsigned row = node / nx, i = node - row * nx;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      gdst[a.pc.lay.at(q, row, i)] = fin[ll.at(q, row, i)];
      gother[a.pc.lay.at(q, row, i)] = before[ll.at(q, row, i)];
    }
    a.rho_out[node] = R[node];
  }
}

// Reduced formats: the same step with the 19 source rows of a 256-node row
// chunk staged through LDS in 16-byte groups, and the results written back
// the same way (through the same LDS buffer), so that each
// global access moves 16 bytes per lane instead of one 2- or 4-byte
// population.  Needs nx and both layout strides to be multiples of the
// group (8 or 4 elements) and a state in pull form (not the post-import
// local form).
constexpr int LDS_GB = 16;  // bytes per staged group
// nodes per block
template <typename S> constexpr int lds_chunk() { return 256; }
// Stage the Q source rows of the chunk [x0, x0 + xc) of row (j, k) -- row
// (j - e_y, k - e_z) of plane q over the window [x0 - G, x0 + XC + G), x
// wrapped periodically -- into win (W = XC + 2G elements per q), one
// 16-byte group per lane and load.  Arguments by value: a kernel argument
// whose address is taken moves to scratch.
template <int D, typename S, int XC>
__device__ inline void stage_rows(Geo g, Layout lay, const S* src, S* win, int x0, int xc, int j, int k) {
  using L = Lat<D>;
  typedef unsigned int V __attribute__((ext_vector_type(4)));  // one 16-byte group
  constexpr int Q = L::Q;
  constexpr int G = LDS_GB / (int)sizeof(S);
  constexpr int W = XC + 2 * G;
  constexpr int NGR = W / G;  // groups per staged row (full chunk)
  const int jm = (j == 0) ? g.ny - 1 : j - 1, jp = (j == g.ny - 1) ? 0 : j + 1;
  int km = 0, kp = 0;
  if (D == 3) { km = (k == 0) ? g.nz - 1 : k - 1; kp = (k == g.nz - 1) ? 0 : k + 1; }
  // signs of e_y, e_z as bit masks over q (no per-lane table lookups)
  constexpr unsigned YP = [] { unsigned m = 0; for (int q = 0; q < Q; q++) m |= (L::ey(q) == 1) << q; return m; }();
  constexpr unsigned YN = [] { unsigned m = 0; for (int q = 0; q < Q; q++) m |= (L::ey(q) == -1) << q; return m; }();
  constexpr unsigned ZP = [] { unsigned m = 0; for (int q = 0; q < Q; q++) m |= (L::ez(q) == 1) << q; return m; }();
  constexpr unsigned ZN = [] { unsigned m = 0; for (int q = 0; q < Q; q++) m |= (L::ez(q) == -1) << q; return m; }();
  const int ngr = (xc == XC) ? NGR : (xc + 2 * G) / G;
  for (int e = threadIdx.x; e < Q * ngr; e += XC) {
    const int q = (xc == XC) ? e / NGR : e / ngr;
    const int gi = e - q * ngr;
    int x = x0 - G + G * gi;
    x = (x < 0) ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    const int ey = (int)((YP >> q) & 1u) - (int)((YN >> q) & 1u);
    const int ez = (int)((ZP >> q) & 1u) - (int)((ZN >> q) & 1u);
    const int sj = ey == 1 ? jm : (ey == -1 ? jp : j);
    const int sk = (D == 3) ? (ez == 1 ? km : (ez == -1 ? kp : k)) : 0;
    const size_t addr = lay.at(q, (unsigned)sj + (unsigned)g.ny * (unsigned)sk, (unsigned)x);
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(src + addr));
    *reinterpret_cast<V*>(&win[q * W + G * gi]) = v;
  }
}

template <int D, typename S, int PM = 0>
__global__ void __launch_bounds__(lds_chunk<S>()) k_step_lds(StepArgs a) {
  constexpr int LDS_XC = lds_chunk<S>();
  using L = Lat<D>;
  using T = typename Store<S>::T;
  typedef unsigned int V __attribute__((ext_vector_type(4)));  // one 16-byte group
  constexpr int Q = L::Q;
  constexpr int G = LDS_GB / (int)sizeof(S);  // elements per group
  constexpr int W = LDS_XC + 2 * G;          // window [x0 - G, x0 + XC + G)
  constexpr int NGO = LDS_XC / G;            // groups per result row (full chunk)
  // one LDS buffer: the staged windows, then (after every lane has pulled
  // its populations into registers) the results
  constexpr int LDS_ELEMS = (Q * W > Q * LDS_XC) ? Q * W : Q * LDS_XC;
  __shared__ __attribute__((aligned(16))) S lds[LDS_ELEMS];
  S (*win)[W] = reinterpret_cast<S (*)[W]>(lds);
  S (*res)[LDS_XC] = reinterpret_cast<S (*)[LDS_XC]>(lds);
  const Geo g = a.pc.g;
  const Layout lay = a.pc.lay;
  const int x0 = blockIdx.x * LDS_XC, j = blockIdx.y, k = blockIdx.z;
  const int xc = (g.nx - x0 < LDS_XC) ? g.nx - x0 : LDS_XC;
  const int t = threadIdx.x;
  const bool live = t < xc;
  const int i = x0 + t;
  const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
  // the IBM force's dependent loads go first; the row staging follows
  T Fp[3] = {0, 0, 0};
  if (PM != 0) {
    int2 run = int2{0, 0};
    if (a.has_ibm) run = pair_run(a, i, (unsigned)j + (unsigned)g.ny * (unsigned)k, live);
    if (live) ibm_node_force<D, S, PM>(a, node, run, Fp[0], Fp[1], Fp[2]);
  }
  stage_rows<D, S, LDS_XC>(g, lay, static_cast<const S*>(a.pc.src), lds, x0, xc, j, k);
  __syncthreads();
  T f[Q];
  if (live)
    pull_node_with<D, S, true>(a.pc, i, j, k, f, [&](int q, int, int, int) {
      return Store<S>::decode(win[q][t + G - L::ex(q)]);
    });
  __syncthreads();
  if (live)
    step_node<D, S, PM>(a, i, j, k, node, f, [&](int q, T v) { res[q][t] = Store<S>::encode(v); },
                        PM != 0 ? Fp : nullptr);
  __syncthreads();
  S* dst = static_cast<S*>(a.dst);
  const unsigned row = (unsigned)j + (unsigned)g.ny * (unsigned)k;
  const int ngo = (xc == LDS_XC) ? NGO : xc / G;
  for (int e = threadIdx.x; e < Q * ngo; e += LDS_XC) {
    const int q = (xc == LDS_XC) ? e / NGO : e / ngo;
    const int gi = e - q * ngo;
    const V v = *reinterpret_cast<const V*>(&res[q][G * gi]);
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(dst + lay.at(q, row, (unsigned)(x0 + G * gi))));
  }
}

// ---------------------------------------------------------------------------
// Macroscopic fields of the current state (compute_macroscopic(false),
// couplb_lattice.h:155-202) with FP64 block partials for the diagnostics
// (couplb_lattice.h:203-242) and the y-profile (fix_couplb.cpp:903-928).
// One block per (row j, plane k); blockDim.x covers nx in chunks.
// partial layout per (j,k): [mass, momx, momy, momz, umax2, clamps,
//                             sum_rho, sum_ux, sum_uy]
// ---------------------------------------------------------------------------
struct MacroArgs {
  PullCtx pc;
  double* rho;  // optional outputs (interior layout), may be null
  double* ux;
  double* uy;
  double* uz;
  double* partial;  // [ny*nz][9]
};

// The nine sums of row (j, k), left in red[c][0] for every thread of the
// block (all threads call it; it ends with a barrier).
template <int D, typename S>
__device__ inline void macro_row(const MacroArgs& a, int j, int k, double (*red)[256]) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const Geo& g = a.pc.g;
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
    T f[Q];
    pull_node<D, S>(a.pc, i, j, k, f);
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
    double r = A + a.pc.base;
    int cl = 0;
    if (r <= RHO_MIN) { r = RHO_MIN; cl = 1; }
    if (!isfinite(r)) { r = RHO_MIN; cl = 1; }
    const double inv_r = 1.0 / r;
    const double u0 = jx * inv_r, u1 = jy * inv_r, u2 = jz * inv_r;
    const unsigned node = (unsigned)i + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
    if (a.rho) { a.rho[node] = r; a.ux[node] = u0; a.uy[node] = u1; a.uz[node] = u2; }
    acc[0] += r;
    acc[1] += r * u0;
    acc[2] += r * u1;
    acc[3] += r * u2;
    const double u2s = u0 * u0 + u1 * u1 + u2 * u2;
    acc[4] = acc[4] > u2s ? acc[4] : u2s;
    acc[5] += cl;
    acc[6] += r;
    acc[7] += u0;
    acc[8] += u1;
  }
  for (int c = 0; c < 9; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      for (int c = 0; c < 9; c++) {
        if (c == 4) red[c][threadIdx.x] = fmax(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        else red[c][threadIdx.x] += red[c][threadIdx.x + s];
      }
    }
    __syncthreads();
  }
}

template <int D, typename S>
__global__ void __launch_bounds__(256) k_macro(MacroArgs a) {
  __shared__ double red[9][256];
  const int j = blockIdx.y, k = blockIdx.z;
  macro_row<D, S>(a, j, k, red);
  if (threadIdx.x < 9) a.partial[((size_t)k * a.pc.g.ny + j) * 9 + threadIdx.x] = red[threadIdx.x][0];
}

// The same sums with a fixed number of partials, MACRO_PARTS, whatever the
// grid size: block t sums the row partials of rows t, t + MACRO_PARTS, ...
// in increasing order, exactly as thread t of k_macro_reduce sums k_macro's
// row partials, so k_macro_reduce over these partials (ny = MACRO_PARTS,
// nz = 1) returns the bits of k_macro + k_macro_reduce.  Every block writes
// its partial (zeros when it has no rows): nothing has to be cleared first.
constexpr int MACRO_PARTS = 256;  // k_macro_reduce's block size
template <int D, typename S>
__global__ void __launch_bounds__(256) k_macro_parts(MacroArgs a) {
  __shared__ double red[9][256];
  const Geo& g = a.pc.g;
  const int nrow = g.ny * (D == 3 ? g.nz : 1);
  double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int r = blockIdx.x; r < nrow; r += MACRO_PARTS) {
    macro_row<D, S>(a, r % g.ny, r / g.ny, red);
    if (threadIdx.x == 0)
      for (int c = 0; c < 9; c++) acc[c] = (c == 4) ? fmax(acc[c], red[c][0]) : acc[c] + red[c][0];
    __syncthreads();  // red is rewritten by the next row
  }
  if (threadIdx.x == 0)
    for (int c = 0; c < 9; c++) a.partial[(size_t)blockIdx.x * 9 + c] = acc[c];
}

// Deterministic second stage: out[0..5] global sums, prof[3*ny] y-profile
// (prof may be null).
// One block; fixed summation order.
__global__ void __launch_bounds__(256) k_macro_reduce(const double* partial, int ny, int nz,
                                                      double* out, double* prof) {
  __shared__ double red[6][256];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  const int nrow = ny * nz;
  for (int r = threadIdx.x; r < nrow; r += blockDim.x) {
    const double* p = partial + (size_t)r * 9;
    for (int c = 0; c < 6; c++) {
      if (c == 4) acc[c] = fmax(acc[c], p[c]);
      else acc[c] += p[c];
    }
  }
  for (int c = 0; c < 6; c++) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 6; c++) {
        if (c == 4) red[c][threadIdx.x] = fmax(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        else red[c][threadIdx.x] += red[c][threadIdx.x + s];
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = red[threadIdx.x][0];
  if (!prof) return;
  for (int j = threadIdx.x; j < ny; j += blockDim.x) {
    double sr = 0, su = 0, sv = 0;
    for (int k = 0; k < nz; k++) {
      const double* p = partial + ((size_t)k * ny + j) * 9;
      sr += p[6]; su += p[7]; sv += p[8];
    }
    prof[3 * j + 0] = sr; prof[3 * j + 1] = su; prof[3 * j + 2] = sv;
  }
}

// ---------------------------------------------------------------------------
// State import/export in the reference's ghost-inclusive layout.
// ---------------------------------------------------------------------------
template <int D, typename S>
__global__ void k_import(const double* fin, void* dstv, Layout lay, Geo g, double base_w_rho0,
                         float scale) {
  using L = Lat<D>;
  constexpr int Q = L::Q;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const int gx = g.nx + 2, gy = g.ny + 2, gz = (D == 3) ? g.nz + 2 : 1;
  const size_t nt = (size_t)gx * gy * gz;
  const size_t gn = (size_t)(i + 1) + (size_t)gx * ((size_t)(j + 1) + (size_t)gy * (D == 3 ? (size_t)(k + 1) : 0));
  const unsigned row = (unsigned)j + (unsigned)g.ny * (unsigned)k;
  S* dst = static_cast<S*>(dstv);
#pragma unroll
  for (int q = 0; q < Q; q++) {
    double v = fin[q * nt + gn];
    if (Store<S>::deviation) v = (v - L::w(q) * base_w_rho0) * scale;
    Store<S>::store(dst + lay.at(q, row, (unsigned)i), (typename Store<S>::T)v);
  }
}

// Grid::init_equilibrium(rho0, ux0, uy0, uz0), couplb_lattice.h:137-154
template <int D, typename S>
__global__ void k_init_eq(void* dstv, Layout lay, Geo g, double r, double u0, double u1,
                          double u2, double base, float scale) {
  using L = Lat<D>;
  constexpr int Q = L::Q;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.nx) return;
  const unsigned row = (unsigned)j + (unsigned)g.ny * (unsigned)k;
  S* dst = static_cast<S*>(dstv);
  const double cs2 = 1.0 / 3.0, cs4 = cs2 * cs2;
  const double uu = u0 * u0 + u1 * u1 + u2 * u2;
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const double eu = L::ex(q) * u0 + L::ey(q) * u1 + L::ez(q) * u2;
    double v = L::w(q) * r * (1.0 + eu / cs2 + 0.5 * eu * eu / cs4 - 0.5 * uu / cs2);
    if (Store<S>::deviation) v = (v - L::w(q) * base) * scale;
    Store<S>::store(dst + lay.at(q, row, (unsigned)i), (typename Store<S>::T)v);
  }
}

__global__ void k_fill(double* p, size_t n, double v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// The reference's f between steps, ghost planes included: interior nodes
// hold the post-stream populations; a ghost node of a periodic face holds
// what Streaming::exchange copied into it before the last stream
// (couplb_streaming.h:136-182, over ghost-inclusive extents in x, y, z
// order) and the swap (couplb_streaming.h:133) moved into f: the
// post-collision populations of the step before the last one at its
// periodic image.  prev is the buffer holding them (the other buffer of the
// pull scheme), or null when there is no such step since the state was set
// (then the images of the current state).  Ghost nodes whose image lies on
// a wall face are never written by the reference after init_equilibrium:
// w_q * rho0.
template <int D, typename S>
__global__ void k_export(PullCtx pc, const void* prev, double* fout) {
  using L = Lat<D>;
  using T = typename Store<S>::T;
  constexpr int Q = L::Q;
  const Geo& g = pc.g;
  const int a = blockIdx.x * blockDim.x + threadIdx.x;  // ghost-inclusive
  const int b = blockIdx.y, c = blockIdx.z;
  const int gx = g.nx + 2, gy = g.ny + 2, gz = (D == 3) ? g.nz + 2 : 1;
  if (a >= gx) return;
  const size_t nt = (size_t)gx * gy * gz;
  const size_t gn = (size_t)a + (size_t)gx * ((size_t)b + (size_t)gy * (size_t)c);
  const int cc = (D == 3) ? c : 1;
  // a ghost coordinate on a periodic face maps to its image; one on a wall
  // face stays a wall ghost
  const bool gxw = (a == 0 && pc.w.type[0][0] > 0) || (a == g.nx + 1 && pc.w.type[0][1] > 0);
  const bool gyw = (b == 0 && pc.w.type[1][0] > 0) || (b == g.ny + 1 && pc.w.type[1][1] > 0);
  const bool gzw = D == 3 && ((cc == 0 && pc.w.type[2][0] > 0) || (cc == g.nz + 1 && pc.w.type[2][1] > 0));
  if (gxw || gyw || gzw) {
#pragma unroll
    for (int q = 0; q < Q; q++) fout[q * nt + gn] = L::w(q) * pc.rho0;
    return;
  }
  const int i = wrap_ghost(a, g.nx), j = wrap_ghost(b, g.ny);
  const int k = (D == 3) ? wrap_ghost(cc, g.nz) : 0;
  const bool ghost = a == 0 || a == g.nx + 1 || b == 0 || b == g.ny + 1 ||
                     (D == 3 && (cc == 0 || cc == g.nz + 1));
  if (ghost && prev) {
    const S* p = static_cast<const S*>(prev);
    const unsigned row = (unsigned)j + (unsigned)g.ny * (unsigned)k;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      T v = Store<S>::load(p + pc.lay.at(q, row, (unsigned)i));
      if (Store<S>::deviation) v *= (T)pc.inv_scale;
      fout[q * nt + gn] = (double)v + L::w(q) * pc.base;
    }
    return;
  }
  T f[Q];
  pull_node<D, S>(pc, i, j, k, f);
#pragma unroll
  for (int q = 0; q < Q; q++) fout[q * nt + gn] = (double)f[q] + L::w(q) * pc.base;
}

}  // namespace couplb
