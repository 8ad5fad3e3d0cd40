// spread_rows.hpp -- force spreading of the IBM records of an LB interval.
//
// IBM::spread (couplb_ibm.h:92-133) adds F_marker * delta * sc into the
// fluid nodes of each marker's 4^D stencil; the self-periodic fold
// (couplb_streaming.h:217-225) then adds the ghost planes into the opposite
// interior faces.  Here every pending record of every sub-step is keyed by
// its base cell, (row of (j0, k0)) * XW + (i0 + 2), and the keys are radix
// sorted once per LB step (rocPRIM); a row table gives each base row's
// range, sorted by i0 within.  One block then gathers the forces of a tile
// of nodes (k_spread_rows): it stages the records that can reach the tile
// into LDS and every node sums their contributions in record order, so the
// spreading is deterministic from run to run and equal to the reference's
// serial sum up to rounding.  Rows of any length and any number of
// sub-steps per LB step are supported.
#pragma once
#include "kernels.hpp"

namespace couplb {

constexpr unsigned REC_SLOT_SHIFT = 24;  // value = slot << 24 | sorted marker index

// width of the i0 range of a base row: i0 + 2 in [0, nx + 4]
__host__ __device__ inline unsigned long long spread_xw(const Geo& g) { return (unsigned long long)g.nx + 5; }

// key of every record of one slot (invalid records get the sentinel and
// sort last)
template <int D>
__global__ void __launch_bounds__(256) k_rec_keys(int n, const MarkerRec* rec, Geo g, unsigned slot,
                                                  unsigned long long sentinel, unsigned long long* key,
                                                  unsigned* val) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const MarkerRec& r = rec[p];
  unsigned long long k = sentinel;
  if (r.valid) {
    const unsigned long long row = (unsigned long long)(r.j0 + 2) +
                                   (unsigned long long)g.nry * (unsigned long long)((D == 3) ? r.k0 + 2 : 0);
    k = row * spread_xw(g) + (unsigned long long)(r.i0 + 2);
  }
  key[p] = k;
  val[p] = (slot << REC_SLOT_SHIFT) | (unsigned)p;
}

// first sorted record of each base row r = 0..nrows (key >= r * XW)
__global__ void __launch_bounds__(256) k_key_rowstart(const unsigned long long* skey, int n, int nrows,
                                                      unsigned long long xw, int* rowstart) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > nrows) return;
  const unsigned long long t = (unsigned long long)r * xw;
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skey[mid] < t) lo = mid + 1; else hi = mid;
  }
  rowstart[r] = lo;
}

struct RowSpreadArgs {
  Geo g;
  Walls w;
  const unsigned long long* key;  // sorted record keys
  const unsigned* val;            // slot << 24 | marker index, in key order
  const int* rowstart;            // per base row, nrows + 1 entries
  const Slot* slots;              // record arrays of the slots
  double* F4;                     // node force (fx, fy, fz, unused)
  unsigned char* flag;            // 1 where some record contributes
  int fold;        // 1: periodic ghost images add into the interior (exchange_forces)
  int accumulate;  // 1: add to the forces already in F4 where flag is set
};

// first sorted index in [lo, hi) whose key is >= t
__device__ inline int lower_key(const unsigned long long* key, int lo, int hi, unsigned long long t) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < t) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Base-cell values of one axis that a tile of ghost-inclusive rows lo..hi
// needs: [lo - 2, hi + 1], plus, when folding, those of the ghost-plane
// images (0 when the tile holds row n and the lo face is periodic: [-2, 1];
// n + 1 when it holds row 1 and the hi face is periodic: [n - 1, n + 2]),
// without repeats.  Single-thread helper; at most hi - lo + 12 values.
__device__ inline int axis_cells(int lo, int hi, int n, int tlo, int thi, int fold, int* out) {
  int c = 0;
  for (int v = lo - 2; v <= hi + 1; v++) out[c++] = v;
  const bool elo = fold && hi >= n && tlo == 0, ehi = fold && lo <= 1 && thi == 0;
  if (elo)
    for (int v = -2; v <= 1; v++)
      if (v < lo - 2 || v > hi + 1) out[c++] = v;
  if (ehi)
    for (int v = n - 1; v <= n + 2; v++)
      if ((v < lo - 2 || v > hi + 1) && !(elo && v >= -2 && v <= 1)) out[c++] = v;
  return c;
}

// Exclusive scan of one int per thread over the block (blockDim.x a multiple
// of 64); total receives the sum.
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

// Tile of the spreading: ST_XS x-positions by ST_TJ x ST_TK rows (3D) or
// by ST_TJ2 rows (2D), one wavefront per z-plane of the tile (3D) or per
// four of its rows (2D), one lane per x-position.
// The records that can reach the tile are those of base rows (j0, k0)
// within two cells (and of the folded ghost images) whose i0 lies in the
// tile's x-windows: per base row one or a few contiguous ranges of the
// sorted keys, staged into LDS window by window.  Every wavefront walks the
// staged records in their order (one record at a time for the whole
// wavefront, read from LDS by broadcast); each lane adds F * delta * sc of
// the records whose stencil covers one of its nodes or their folded ghost
// images, so each node sums its contributions in record order:
// deterministic, and the reference's sum up to rounding.
constexpr int ST_THREADS = 256;
constexpr int ST_XS = 64;
constexpr int ST_TJ = 4, ST_TK = 4;  // 3D tile rows
constexpr int ST_TJ2 = 16;           // 2D tile rows
constexpr int ST_AXIS = 32;          // base-cell values per axis
constexpr int ST_PAIRS = 704;        // (base row, x-window) ranges per tile (<= 15 * 15 * 3)
constexpr int ST_WIN = 64;           // staged records per LDS window
constexpr int ST_NV = 16;            // doubles staged per record

template <int D>
__global__ void __launch_bounds__(ST_THREADS) k_spread_rows(RowSpreadArgs a) {
#pragma clang fp contract(off)
  __shared__ int Yl[ST_AXIS], Zl[ST_AXIS], nyl, nzl, nwin, wl[3], wh[3];
  __shared__ int pst[ST_PAIRS], poff[ST_PAIRS + 1], sh[ST_THREADS / 64];
  __shared__ int ci0[ST_WIN], cj0[ST_WIN], ck0[ST_WIN];
  __shared__ double cv[ST_NV][ST_WIN];
  const Geo& g = a.g;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int x0 = blockIdx.x * ST_XS;
  const int xl = (g.nx - x0 < ST_XS) ? g.nx - x0 : ST_XS;
  const int TJ = (D == 3) ? ST_TJ : ST_TJ2;
  const int J0 = blockIdx.y * TJ, K0 = (D == 3) ? blockIdx.z * ST_TK : 0;
  // this wavefront's rows: jw .. jw + 3 of plane kw
  const int jw = (D == 3) ? J0 : J0 + 4 * wv;
  const int kw = (D == 3) ? K0 + wv : 0;
  if (tid == 0) {
    const int jh = (J0 + TJ < g.ny ? J0 + TJ : g.ny);
    nyl = axis_cells(J0 + 1, jh, g.ny, a.w.type[1][0], a.w.type[1][1], a.fold, Yl);
    if (D == 3) {
      const int kh = (K0 + ST_TK < g.nz ? K0 + ST_TK : g.nz);
      nzl = axis_cells(K0 + 1, kh, g.nz, a.w.type[2][0], a.w.type[2][1], a.fold, Zl);
    } else {
      nzl = 1;
      Zl[0] = 0;
    }
    // x-windows of i0: the tile's own, and those of folded ghost images,
    // merged so that no record is staged twice
    int lo[3], hi[3], m = 0;
    lo[m] = x0 - 1; hi[m] = x0 + xl + 1; m++;
    if (a.fold && a.w.type[0][0] == 0 && x0 + xl == g.nx) { lo[m] = -2; hi[m] = 1; m++; }
    if (a.fold && a.w.type[0][1] == 0 && x0 == 0) { lo[m] = g.nx - 1; hi[m] = g.nx + 2; m++; }
    for (int p = 1; p < m; p++)
      for (int q = p; q > 0 && lo[q] < lo[q - 1]; q--) {
        const int tl = lo[q], th = hi[q];
        lo[q] = lo[q - 1]; hi[q] = hi[q - 1];
        lo[q - 1] = tl; hi[q - 1] = th;
      }
    int nw = 0;
    for (int p = 0; p < m; p++) {
      if (nw > 0 && lo[p] <= wh[nw - 1] + 1) {
        wh[nw - 1] = hi[p] > wh[nw - 1] ? hi[p] : wh[nw - 1];
      } else {
        wl[nw] = lo[p]; wh[nw] = hi[p]; nw++;
      }
    }
    nwin = nw;
  }
  __syncthreads();
  const unsigned long long xw = spread_xw(g);
  const int nY = nyl, nZ = nzl, nw = nwin;
  const int npair = nY * nZ * nw;  // pair = (zi * nY + yi) * nw + w
  // ranges of every (base row, window): first key and count; offsets by scan
  int carry = 0;
  for (int p0 = 0; p0 < npair; p0 += ST_THREADS) {
    const int p = p0 + tid;
    int cnt = 0;
    if (p < npair) {
      const int w = p % nw, bi = p / nw;
      const int yi = bi % nY, zi = bi / nY;
      const int row = (Yl[yi] + 2) + g.nry * ((D == 3) ? Zl[zi] + 2 : 0);
      const int r0 = a.rowstart[row], r1 = a.rowstart[row + 1];
      const unsigned long long rk = (unsigned long long)row * xw;
      const unsigned long long klo = rk + (unsigned long long)(wl[w] + 2);
      const unsigned long long khi = rk + (unsigned long long)(wh[w] + 3);
      int e0, e1;
      if (r1 - r0 <= 8) {
        // short row: its keys in one round of loads, bounds counted
        unsigned long long kk[8];
#pragma unroll
        for (int m = 0; m < 8; m++) kk[m] = (r0 + m < r1) ? a.key[r0 + m] : ~0ull;
        e0 = r0;
        e1 = r0;
#pragma unroll
        for (int m = 0; m < 8; m++) {
          e0 += (kk[m] < klo) ? 1 : 0;
          e1 += (kk[m] < khi) ? 1 : 0;
        }
      } else {
        e0 = lower_key(a.key, r0, r1, klo);
        e1 = lower_key(a.key, e0, r1, khi);
      }
      pst[p] = e0;
      cnt = e1 - e0;
    }
    int tot;
    const int off = block_exscan(cnt, sh, tot);
    if (p < npair) poff[p] = carry + off;
    carry += tot;
  }
  if (tid == 0) poff[npair] = carry;
  __syncthreads();
  const int total = carry;
  // this lane's node x and its images; the wavefront's planes / rows
  const int x = x0 + lane;
  const bool xin = lane < xl;
  const Img xa = (a.fold && xin) ? images(x + 1, g.nx, a.w.type[0][0], a.w.type[0][1]) : Img{1, x + 1, -9, -9};
  const Img za = (D == 3) ? (a.fold ? images(kw + 1, g.nz, a.w.type[2][0], a.w.type[2][1]) : Img{1, kw + 1, -9, -9})
                          : Img{1, 0, -9, -9};
  const bool kin = (D == 3) ? kw < g.nz : true;
  Img ya[4];
#pragma unroll
  for (int r = 0; r < 4; r++)
    ya[r] = (a.fold && jw + r < g.ny) ? images(jw + r + 1, g.ny, a.w.type[1][0], a.w.type[1][1])
                                       : Img{jw + r < g.ny ? 1 : 0, jw + r + 1, -9, -9};
  double acc[4][3];
  bool hit[4];
#pragma unroll
  for (int r = 0; r < 4; r++) { acc[r][0] = acc[r][1] = acc[r][2] = 0.0; hit[r] = false; }

  for (int w0 = 0; w0 < total; w0 += ST_WIN) {
    const int nst = (total - w0 < ST_WIN) ? total - w0 : ST_WIN;
    for (int q = tid; q < nst; q += ST_THREADS) {
      const int gq = w0 + q;
      int lo = 0, hi = npair - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (poff[mid] <= gq) lo = mid; else hi = mid - 1;
      }
      const int e = pst[lo] + (gq - poff[lo]);
      const int bi = lo / nw;
      const int j0 = Yl[bi % nY], k0 = (D == 3) ? Zl[bi / nY] : 0;
      const int row = (j0 + 2) + g.nry * ((D == 3) ? k0 + 2 : 0);
      ci0[q] = (int)(a.key[e] - (unsigned long long)row * xw) - 2;
      cj0[q] = j0;
      ck0[q] = k0;
      const unsigned vv = a.val[e];
      // the record's first 16 doubles (wx, wy, wz, sc, fx, fy, fz) in eight
      // 16-byte loads (records are 144 bytes, so 16-byte aligned)
      const double2* rp = reinterpret_cast<const double2*>(
          a.slots[vv >> REC_SLOT_SHIFT].rec + (vv & ((1u << REC_SLOT_SHIFT) - 1u)));
      double2 rv[ST_NV / 2];
#pragma unroll
      for (int m = 0; m < ST_NV / 2; m++) rv[m] = rp[m];
#pragma unroll
      for (int m = 0; m < ST_NV / 2; m++) {
        cv[2 * m][q] = rv[m].x;
        cv[2 * m + 1][q] = rv[m].y;
      }
    }
    __syncthreads();
    if (kin) {
      // every lane walks the staged records in order (the same record for
      // the whole wavefront) and adds those whose stencil covers its nodes
      for (int e = 0; e < nst; e++) {
        const int i0 = ci0[e], j0 = cj0[e], k0 = ck0[e];
        each_image(za, [&](int c) {
          const int v = (D == 3) ? c - k0 + 1 : 0;
          if (v < 0 || v > 3) return;
          const double wzv = (D == 3) ? cv[8 + v][e] : 1.0;
          const double sc = cv[12][e], fx = cv[13][e], fy = cv[14][e], fz = cv[15][e];
#pragma unroll
          for (int r = 0; r < 4; r++) {
            each_image(ya[r], [&](int b) {
              const int u = b - j0 + 1;
              if (u < 0 || u > 3) return;
              const double wyu = cv[4 + u][e];
              each_image(xa, [&](int aa) {
                const int t = aa - i0 + 1;
                if (!xin || t < 0 || t > 3) return;
                const double d = (D == 3) ? (cv[t][e] * wyu) * wzv : cv[t][e] * wyu;
                const double wgt = d * sc;
                acc[r][0] += fx * wgt;
                acc[r][1] += fy * wgt;
                acc[r][2] += fz * wgt;
                hit[r] = true;
              });
            });
          }
        });
      }
    }
    __syncthreads();  // the next window overwrites the staged records
  }
  if (!kin || !xin) return;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int j = jw + r;
    if (j >= g.ny) continue;
    const size_t node = (size_t)x + (size_t)g.nx * ((size_t)j + (size_t)g.ny * (size_t)kw);
    if (a.accumulate) {
      if (hit[r]) {
        double4* F = reinterpret_cast<double4*>(a.F4 + 4 * node);
        if (a.flag[node]) {
          const double4 o = *F;
          *F = double4{o.x + acc[r][0], o.y + acc[r][1], o.z + acc[r][2], 0.0};
        } else {
          *F = double4{acc[r][0], acc[r][1], acc[r][2], 0.0};
        }
        a.flag[node] = 1;
      }
    } else {
      a.flag[node] = hit[r] ? 1 : 0;
      if (hit[r]) *reinterpret_cast<double4*>(a.F4 + 4 * node) = double4{acc[r][0], acc[r][1], acc[r][2], 0.0};
    }
  }
}

// dense interior node forces from the spreading's output: out = F where
// flagged, else 0
__global__ void __launch_bounds__(256) k_force_export(size_t N, const double* F4,
                                                     const unsigned char* flag, double* out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const bool h = flag[i] != 0;
#pragma unroll
  for (int c = 0; c < 3; c++) out[3 * i + c] = h ? F4[4 * i + c] : 0.0;
}

}  // namespace couplb
