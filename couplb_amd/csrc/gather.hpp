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
  int TJ, TK, XM;  // rows in y and z per tile, x positions per thread
};

// Tile variants: a thread owns XM x-positions of each of the TJ*TK rows
// (at most NPT nodes), so a tile holds at most TILE_NODES nodes.
__host__ __device__ inline TileShape tile_shape(int D, int nx) {
  if (nx <= TILE_THREADS) return D == 3 ? TileShape{2, 2, 1} : TileShape{4, 1, 1};
  if (nx <= 2 * TILE_THREADS) return TileShape{2, 1, 2};
  return TileShape{1, 1, 4};
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
// staged window.  All threads of the block call this together.  Candidate
// q of the tile (rows in slot, z, y order, each row's records in order)
// lands in window entry q - w0; one thread per candidate loads it, finding
// its row by binary search over the rows' offsets, so the loads of a window
// are all in flight together.
template <int D, typename Slots, typename Cand, int WIN, typename Load, typename Use>
__device__ inline void stage_tile(const Slots& slots, int nslots, const Geo& g, const TileLists& tl,
                                  Cand* win, int* sh, Load&& load, Use&& use) {
  __shared__ int row_off[TILE_THREADS], row_r0[TILE_THREADS], row_slot[TILE_THREADS];
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
    row_off[threadIdx.x] = off;
    row_r0[threadIdx.x] = r0;
    row_slot[threadIdx.x] = s;
    __syncthreads();
    const int nrow = min((int)blockDim.x, R - g0);
    for (int w0 = 0; w0 < total; w0 += WIN) {
      const int nw = min(WIN, total - w0);
      for (int q = (int)threadIdx.x; q < nw; q += blockDim.x) {
        const int gq = w0 + q;
        // the last row starting at or before gq holds it (rows without
        // records share their offset with the next row)
        int lo = 0, hi = nrow - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (row_off[mid] <= gq) lo = mid;
          else hi = mid - 1;
        }
        load(row_slot[lo], row_r0[lo] + (gq - row_off[lo]), win[q]);
      }
      __syncthreads();
      use(nw);
      __syncthreads();
    }
    __syncthreads();  // row_* are rewritten by the next group of rows
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

constexpr int SPREAD_WIN = 160;  // MarkerRec per LDS window (22.5 KiB)

struct RecSlots {
  const Slot* sl;
  __device__ const int* rowstart(int s) const { return sl[s].rowstart; }
};

struct SpreadArgs {
  Geo g;
  Walls w;
  int nslots;
  const Slot* slots;    // device table of nslots descriptors
  double* F4;  // node force (fx, fy, fz, unused), 32 B per node
  unsigned char* flag;  // per node: 1 where some record contributes
  int fold;  // 1: periodic ghost images add into the interior (exchange_forces);
             // 0: direct contributions only (before the interval's last sub-step)
};

// Per-thread view of its tile nodes: XM x-positions x TJ*TK rows, with the
// ghost-plane images of every coordinate computed once (fold = false: the
// node itself only).
template <int D, int TJ, int TK, int XM>
struct TileNodes {
  int x[XM], j[TJ], k[TK];
  Img xa[XM], ya[TJ], za[TK];
  bool xin[XM], yin[TJ], zin[TK];
  __device__ TileNodes(const Geo& g, const Walls& w, int J0, int K0, bool fold = true) {
#pragma unroll
    for (int m = 0; m < XM; m++) {
      x[m] = threadIdx.x + m * TILE_THREADS;
      xin[m] = x[m] < g.nx;
      xa[m] = fold ? images(x[m] + 1, g.nx, w.type[0][0], w.type[0][1]) : Img{1, x[m] + 1, -9, -9};
    }
#pragma unroll
    for (int r = 0; r < TJ; r++) {
      j[r] = J0 - 1 + r;
      yin[r] = j[r] < g.ny;
      ya[r] = fold ? images(j[r] + 1, g.ny, w.type[1][0], w.type[1][1]) : Img{1, j[r] + 1, -9, -9};
    }
#pragma unroll
    for (int r = 0; r < TK; r++) {
      k[r] = K0 - 1 + r;
      zin[r] = (D == 3) ? k[r] < g.nz : r == 0;
      za[r] = (D == 3) ? (fold ? images(k[r] + 1, g.nz, w.type[2][0], w.type[2][1])
                               : Img{1, k[r] + 1, -9, -9})
                       : Img{1, 0, -9, -9};
    }
  }
};

// IBM::spread for every record of the pending slots, as a per-node gather
// (couplb_ibm.h:92-133; ghost-plane images carry the self-periodic fold of
// couplb_streaming.h:217-225).  One block per tile of rows.  Per node the
// contributions are summed in the fixed order (window, x-image, base cell,
// record, y-image, z-image).
template <int D, int TJ, int TK, int XM>
__global__ void __launch_bounds__(TILE_THREADS) k_spread(SpreadArgs a) {
  __shared__ MarkerRec win[SPREAD_WIN];
  __shared__ int cnt[MAX_BIN], start[MAX_BIN + 1], list[SPREAD_WIN], sh[TILE_THREADS / 64];
  __shared__ TileLists tl;
  const Geo g = a.g;
  const Walls w = a.w;
  const TileShape ts{TJ, TK, XM};
  const int J0 = 1 + blockIdx.y * TJ, K0 = 1 + blockIdx.z * TK;
  const int nbin = g.nx + 5;
  tile_lists<D>(g, w, J0, K0, ts, tl);
  const TileNodes<D, TJ, TK, XM> tn(g, w, J0, K0, a.fold != 0);
  double acc[XM][TJ][TK][3];
  bool hit[XM][TJ][TK];
#pragma unroll
  for (int m = 0; m < XM; m++)
#pragma unroll
    for (int r = 0; r < TJ; r++)
#pragma unroll
      for (int q = 0; q < TK; q++) {
        acc[m][r][q][0] = acc[m][r][q][1] = acc[m][r][q][2] = 0.0;
        hit[m][r][q] = false;
      }
  RecSlots slots{a.slots};
  stage_tile<D, RecSlots, MarkerRec, SPREAD_WIN>(
      slots, a.nslots, g, tl, win, sh,
      [&](int s, int idx, MarkerRec& c) { c = a.slots[s].rec[idx]; },
      [&](int n) {
        build_bins(n, nbin, cnt, start, list, sh,
                   [&](int c) { return win[c].valid ? win[c].i0 + 2 : -1; });
#pragma unroll
        for (int m = 0; m < XM; m++) {
#pragma clang fp contract(off)
          if (!tn.xin[m]) continue;
          each_image(tn.xa[m], [&](int aa) {
            for (int kb = aa; kb <= aa + 3 && kb < nbin; kb++) {  // base cells aa-2 .. aa+1
              const int t = aa - kb + 3;
              for (int e = start[kb]; e < start[kb + 1]; e++) {
                const MarkerRec& c = win[list[e]];
                const double wxt = c.wx[t];
#pragma unroll
                for (int r = 0; r < TJ; r++) {
                  if (!tn.yin[r]) continue;
#pragma unroll
                  for (int q = 0; q < TK; q++) {
                    if (!tn.zin[q]) continue;
                    each_image(tn.ya[r], [&](int bb) {
                      const int u = bb - c.j0 + 1;
                      if (u < 0 || u > 3) return;
                      each_image(tn.za[q], [&](int cc) {
                        const int v = (D == 3) ? cc - c.k0 + 1 : 0;
                        if (v < 0 || v > 3) return;
                        const double d = (D == 3) ? (wxt * c.wy[u]) * c.wz[v] : wxt * c.wy[u];
                        const double wgt = d * c.sc;
                        acc[m][r][q][0] += c.fx * wgt;
                        acc[m][r][q][1] += c.fy * wgt;
                        acc[m][r][q][2] += c.fz * wgt;
                        hit[m][r][q] = true;
                      });
                    });
                  }
                }
              }
            }
          });
        }
      });
#pragma unroll
  for (int m = 0; m < XM; m++)
#pragma unroll
    for (int r = 0; r < TJ; r++)
#pragma unroll
      for (int q = 0; q < TK; q++) {
        if (!tn.xin[m] || !tn.yin[r] || !tn.zin[q]) continue;
        const unsigned node = (unsigned)tn.x[m] + (unsigned)g.nx * ((unsigned)tn.j[r] + (unsigned)g.ny * (unsigned)tn.k[q]);
        a.flag[node] = hit[m][r][q] ? 1 : 0;
        if (hit[m][r][q]) {
          *reinterpret_cast<double4*>(a.F4 + 4 * (size_t)node) =
              double4{acc[m][r][q][0], acc[m][r][q][1], acc[m][r][q][2], 0.0};
        }
      }
}

// dense interior node forces from k_spread's output: out = F where flagged, else 0
__global__ void __launch_bounds__(256) k_force_export(size_t N, const double* F4,
                                                     const unsigned char* flag, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool h = flag[i] != 0;
#pragma unroll
  for (int c = 0; c < 3; c++) out[3 * i + c] = h ? F4[4 * i + c] : 0.0;
}

}  // namespace couplb
