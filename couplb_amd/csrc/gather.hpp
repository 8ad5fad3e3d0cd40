// gather.hpp -- tiled candidate staging for the IBM gathers.
//
// Markers are kept sorted by the row (j0, k0) of their base cell, with a
// row table giving each row's range.  One block owns a tile of TJ x TK
// interior rows.  It needs every record whose 4^D stencil reaches an image of
// one of its nodes: base rows j0 within two cells of the tile's y-rows (and
// of their ghost-plane images), likewise k0.  Those base rows are listed once
// per tile; their ranges (slot, k0, j0) are counted by one thread each,
// prefix-summed across the block and copied cooperatively into LDS, window
// by window.  Each window is binned by base x-cell (a counting sort in LDS
// whose bins are put in candidate order), so the node at x visits only the
// bins of base cells x-2..x+1.  Work is proportional to the contributions,
// the loads are parallel, and the visiting order is fixed: every reduction
// built on it is deterministic.
#pragma once
#include "kernels.hpp"

namespace couplb {

constexpr int TILE_THREADS = 256;
constexpr int NPT = 4;                           // node slots per thread
constexpr int TILE_NODES = TILE_THREADS * NPT;   // nodes per tile
constexpr int MAX_NX_IBM = TILE_NODES;           // rows up to 1024 nodes
constexpr int MAX_BIN = MAX_NX_IBM + 5;          // base x-cells -2 .. nx+2
constexpr int MAX_AXIS_ROWS = 32;                // listed base rows per axis

struct TileShape {
  int TJ, TK;
};

// Rows per tile so a tile holds at most TILE_NODES nodes.
__host__ __device__ inline TileShape tile_shape(int D, int nx) {
  if (D == 2) {
    int tj = TILE_NODES / (nx > 0 ? nx : 1);
    tj = tj > 16 ? 16 : (tj < 1 ? 1 : tj);
    return {tj, 1};
  }
  if (nx <= TILE_NODES / 4) return {2, 2};
  if (nx <= TILE_NODES / 2) return {2, 1};
  return {1, 1};
}

// Base-row values of one axis a tile covering ghost-inclusive rows lo..hi
// needs: [lo-2, hi+1], plus the windows of the ghost-plane images (0 when
// the tile holds row n and the lo face is periodic; n+1 when it holds row 1
// and the hi face is periodic), without repeats.  Single-thread helper.
__device__ inline int axis_rows(int lo, int hi, int n, int tlo, int thi, int* out) {
  int c = 0;
  for (int v = lo - 2; v <= hi + 1; v++) out[c++] = v;
  const bool elo = (hi >= n && tlo == 0), ehi = (lo <= 1 && thi == 0);
  if (elo)
    for (int v = -2; v <= 1; v++)
      if (v < lo - 2 || v > hi + 1) out[c++] = v;
  if (ehi)
    for (int v = n - 1; v <= n + 2; v++)
      if ((v < lo - 2 || v > hi + 1) && !(elo && v >= -2 && v <= 1)) out[c++] = v;
  return c;
}

// Exclusive block scan of one int per thread (blockDim.x a multiple of 64).
__device__ inline int block_exscan(int x, int* sh, int& total) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(v, o);
    if (lane >= o) v += y;
  }
  if (lane == 63) sh[wid] = v;
  __syncthreads();
  int base = 0, tot = 0;
  for (int w = 0; w < nw; w++) {
    const int s = sh[w];
    if (w < wid) base += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return base + v - x;
}

struct TileLists {
  int Y[MAX_AXIS_ROWS], Z[MAX_AXIS_ROWS];
  int nY, nZ;
};

// List the tile's base rows (one thread) and share them through LDS.
template <int D>
__device__ inline void tile_lists(const Geo& g, const Walls& w, int J0, int K0, TileShape ts,
                                  TileLists& tl) {
  if (threadIdx.x == 0) {
    const int Jh = min(J0 + ts.TJ - 1, g.ny);
    tl.nY = axis_rows(J0, Jh, g.ny, w.type[1][0], w.type[1][1], tl.Y);
    if (D == 3) {
      const int Kh = min(K0 + ts.TK - 1, g.nz);
      tl.nZ = axis_rows(K0, Kh, g.nz, w.type[2][0], w.type[2][1], tl.Z);
    } else {
      tl.nZ = 1;
      tl.Z[0] = 0;
    }
  }
  __syncthreads();
}

// Walk every candidate of the tile window by window.  load(slot, idx, dst)
// copies record idx of the slot into LDS entry dst; use(count) consumes the
// staged window.  All threads of the block call this together.
template <int D, typename Slots, typename Cand, int WIN, typename Load, typename Use>
__device__ inline void stage_tile(const Slots& slots, int nslots, const Geo& g, const TileLists& tl,
                                  Cand* win, int* sh, Load&& load, Use&& use) {
  const int per = tl.nY * tl.nZ;
  const int R = nslots * per;
  for (int g0 = 0; g0 < R; g0 += blockDim.x) {
    const int r = g0 + (int)threadIdx.x;
    int s = 0, r0 = 0, cnt = 0;
    if (r < R) {
      s = r / per;
      const int rem = r - s * per;
      const int zi = rem / tl.nY, yi = rem - zi * tl.nY;
      const int row = (tl.Y[yi] + 2) + g.nry * ((D == 3) ? (tl.Z[zi] + 2) : 0);
      r0 = slots.rowstart(s)[row];
      cnt = slots.rowstart(s)[row + 1] - r0;
    }
    int total;
    const int off = block_exscan(cnt, sh, total);
    for (int w0 = 0; w0 < total; w0 += WIN) {
      const int lo = off > w0 ? off : w0;
      const int hi = (off + cnt) < (w0 + WIN) ? (off + cnt) : (w0 + WIN);
      for (int q = lo; q < hi; q++) load(s, r0 + (q - off), win[q - w0]);
      __syncthreads();
      use(total - w0 < WIN ? total - w0 : WIN);
      __syncthreads();
    }
  }
}

// Counting sort of window entries 0..n-1 by key(c) in [0, nbin) (negative:
// skip): start[b]..start[b+1] lists the entries of bin b, in entry order.
template <typename Key>
__device__ inline void build_bins(int n, int nbin, int* cnt, int* start, int* list, int* sh,
                                  Key&& key) {
  const int t = threadIdx.x, bd = blockDim.x;
  for (int b = t; b < nbin; b += bd) cnt[b] = 0;
  __syncthreads();
  for (int c = t; c < n; c += bd) {
    const int k = key(c);
    if (k >= 0 && k < nbin) atomicAdd(&cnt[k], 1);
  }
  __syncthreads();
  int carry = 0;
  for (int b0 = 0; b0 < nbin; b0 += bd) {
    const int b = b0 + t;
    const int v = b < nbin ? cnt[b] : 0;
    int tot;
    const int ex = block_exscan(v, sh, tot);
    if (b < nbin) start[b] = carry + ex;
    carry += tot;
  }
  if (t == 0) start[nbin] = carry;
  for (int b = t; b < nbin; b += bd) cnt[b] = 0;
  __syncthreads();
  for (int c = t; c < n; c += bd) {
    const int k = key(c);
    if (k >= 0 && k < nbin) list[start[k] + atomicAdd(&cnt[k], 1)] = c;
  }
  __syncthreads();
  for (int b = t; b < nbin; b += bd) {  // entry order inside each bin
    const int lo = start[b], hi = start[b + 1];
    for (int e = lo + 1; e < hi; e++) {
      const int x = list[e];
      int f = e - 1;
      while (f >= lo && list[f] > x) { list[f + 1] = list[f]; f--; }
      list[f + 1] = x;
    }
  }
  __syncthreads();
}

// node slot p of this thread -> interior coordinates (x, j, k); false if the
// slot falls outside the (clipped) tile
__device__ inline bool tile_node(const Geo& g, TileShape ts, int J0, int K0, int p, int& x, int& j,
                                 int& k) {
  const int idx = threadIdx.x + p * blockDim.x;
  if (idx >= ts.TJ * ts.TK * g.nx) return false;
  x = idx % g.nx;
  const int r = idx / g.nx;
  j = J0 - 1 + r % ts.TJ;
  k = K0 - 1 + r / ts.TJ;
  return j < g.ny && k < (g.D == 3 ? g.nz : 1);
}

constexpr int SPREAD_WIN = 96;   // MarkerRec per LDS window (13.5 KiB)
constexpr int COVER_WIN = 512;   // base cells per LDS window (8 KiB)

struct RecSlots {
  const Slot* sl;
  __device__ const int* rowstart(int s) const { return sl[s].rowstart; }
};

struct SpreadArgs {
  Geo g;
  Walls w;
  int nslots;
  const Slot* slots;    // device table of nslots descriptors
  double* Fx;
  double* Fy;
  double* Fz;
  unsigned char* flag;  // per node: 1 where some record contributes
};

// IBM::spread for every record of the pending slots, as a per-node gather
// (couplb_ibm.h:92-133; ghost-plane images carry the self-periodic fold of
// couplb_streaming.h:217-225).  One block per tile of rows.
template <int D>
__global__ void __launch_bounds__(TILE_THREADS) k_spread(SpreadArgs a) {
  __shared__ MarkerRec win[SPREAD_WIN];
  __shared__ int cnt[MAX_BIN], start[MAX_BIN + 1], list[SPREAD_WIN], sh[TILE_THREADS / 64];
  __shared__ TileLists tl;
  const Geo g = a.g;
  const Walls w = a.w;
  const TileShape ts = tile_shape(D, g.nx);
  const int J0 = 1 + blockIdx.y * ts.TJ, K0 = 1 + blockIdx.z * ts.TK;
  const int nbin = g.nx + 5;
  tile_lists<D>(g, w, J0, K0, ts, tl);
  double ax[NPT], ay[NPT], az[NPT];
  bool hit[NPT];
#pragma unroll
  for (int p = 0; p < NPT; p++) { ax[p] = ay[p] = az[p] = 0.0; hit[p] = false; }
  RecSlots slots{a.slots};
  stage_tile<D, RecSlots, MarkerRec, SPREAD_WIN>(
      slots, a.nslots, g, tl, win, sh,
      [&](int s, int idx, MarkerRec& c) { c = a.slots[s].rec[idx]; },
      [&](int n) {
        build_bins(n, nbin, cnt, start, list, sh,
                   [&](int c) { return win[c].valid ? win[c].i0 + 2 : -1; });
#pragma unroll
        for (int p = 0; p < NPT; p++) {
#pragma clang fp contract(off)
          int x, j, k;
          if (!tile_node(g, ts, J0, K0, p, x, j, k)) continue;
          const Img xa = images(x + 1, g.nx, w.type[0][0], w.type[0][1]);
          const Img ya = images(j + 1, g.ny, w.type[1][0], w.type[1][1]);
          const Img za = (D == 3) ? images(k + 1, g.nz, w.type[2][0], w.type[2][1]) : Img{1, 0, -9, -9};
          for (int ia = 0; ia < xa.n; ia++) {
            const int aa = xa.at(ia);
            for (int ib = 0; ib < ya.n; ib++) {
              const int bb = ya.at(ib);
              for (int ic = 0; ic < za.n; ic++) {
                const int cc = za.at(ic);
                for (int kb = aa; kb <= aa + 3; kb++) {  // base cells aa-2 .. aa+1
                  if (kb >= nbin) break;
                  const int t = aa - kb + 3;
                  for (int e = start[kb]; e < start[kb + 1]; e++) {
                    const MarkerRec& m = win[list[e]];
                    const int u = bb - m.j0 + 1;
                    const int v = (D == 3) ? cc - m.k0 + 1 : 0;
                    if (u < 0 || u > 3 || v < 0 || v > 3) continue;
                    const double d = (D == 3) ? (m.wx[t] * m.wy[u]) * m.wz[v] : m.wx[t] * m.wy[u];
                    const double wgt = d * m.sc;
                    ax[p] += m.fx * wgt;
                    ay[p] += m.fy * wgt;
                    az[p] += m.fz * wgt;
                    hit[p] = true;
                  }
                }
              }
            }
          }
        }
      });
#pragma unroll
  for (int p = 0; p < NPT; p++) {
    int x, j, k;
    if (!tile_node(g, ts, J0, K0, p, x, j, k)) continue;
    const unsigned node = (unsigned)x + (unsigned)g.nx * ((unsigned)j + (unsigned)g.ny * (unsigned)k);
    a.flag[node] = hit[p] ? 1 : 0;
    if (hit[p]) {
      a.Fx[node] = ax[p];
      a.Fy[node] = ay[p];
      a.Fz[node] = az[p];
    }
  }
}

struct BaseSlots {
  const int* rs;
  __device__ const int* rowstart(int) const { return rs; }
};

}  // namespace couplb
