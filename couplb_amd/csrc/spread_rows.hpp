// spread_rows.hpp -- force spreading of the IBM records of an LB interval.
//
// IBM::spread (couplb_ibm.h:92-133) adds F_marker * delta * sc into the
// fluid nodes of each marker's 4^D stencil; the self-periodic fold
// (couplb_streaming.h:217-225) then adds the ghost planes into the opposite
// interior faces.
//
// Here the pending records of every sub-step are binned once per LB step by
// a counting sort: the bucket of a record is its base row (j0, k0) and a
// 16-cell segment of its i0, the bucket counts are scanned into bucket
// starts, every record is placed into its bucket, and each bucket is sorted
// by (slot, marker), so the order of the records is a pure function of the
// records.  One wavefront then owns a tile of nodes (k_spread_tiles): it
// stages the records that can reach the tile, and walks them in that order
// with one lane per stencil node of the record (4 x 4 x 4 lanes in 3-D),
// each lane adding F * delta * sc into the node's accumulator in LDS.  The
// 64 stencil nodes of one record are 64 different nodes, so every node sums
// its contributions in record order: deterministic from run to run and equal
// to the reference's serial sum up to rounding.  Rows of any length and any
// number of sub-steps per LB step are supported.  The same kernels spread
// only the records that left their ordered base cell (moving markers, whose
// other records go through the moving-marker map of spread_pairs.hpp): then
// only the tiles those records reach are listed and spread.
#pragma once
#include "kernels.hpp"

namespace couplb {

constexpr unsigned REC_SLOT_SHIFT = 24;  // entry value = slot << 24 | sorted marker index
constexpr int SB_X = 16;                 // i0 cells per bucket

// buckets per base row: i0 + 2 in [0, nx + 4]
__host__ __device__ inline int spread_nxb(const Geo& g) { return (g.nx + 5 + SB_X - 1) / SB_X; }

__global__ void __launch_bounds__(256) k_fill_int(int* p, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Wavefront tile of the spreading, in interior nodes, and the pitches of
// its LDS accumulators: a row pitch of 20 doubles and a plane pitch of 80
// put the 4 x 4 x 4 lanes of one record on different LDS banks.
template <int D> struct TileShape;
template <> struct TileShape<3> { static constexpr int X = 16, Y = 4, Z = 4, PX = 20, PY = 4; };
template <> struct TileShape<2> { static constexpr int X = 16, Y = 16, Z = 1, PX = 20, PY = 16; };

// Interior node (ghost-inclusive coordinate 1..n) that ghost-inclusive
// stencil coordinate a stands for on an axis of n nodes, or 0 if none:
// the node itself, or across a periodic face the node whose ghost image it
// is (ghost 0 is node n, ghost n + 1 is node 1); stencil nodes outside
// [0, n + 1] are dropped (the reference's one-ghost truncation).
__device__ inline int axis_node(int a, int n, bool plo, bool phi, int fold) {
  if (a >= 1 && a <= n) return a;
  if (fold && a == 0 && plo) return n;
  if (fold && a == n + 1 && phi) return 1;
  return 0;
}

// bucket of a record from its base cell and validity (i0, j0, k0, valid);
// -1 if it contributes nothing (invalid, or a base cell whose stencil misses
// the ghost-inclusive grid)
template <int D>
__device__ inline int rec_bucket(const int4& r, const Geo& g) {
  if (!r.w) return -1;
  if (r.x < -2 || r.x > g.nx + 2 || r.y < -2 || r.y > g.ny + 2) return -1;
  if (D == 3 && (r.z < -2 || r.z > g.nz + 2)) return -1;
  const int row = (r.y + 2) + g.nry * ((D == 3) ? r.z + 2 : 0);
  return row * spread_nxb(g) + (r.x + 2) / SB_X;
}

// base cell and validity of a record
__device__ inline int4 rec_cell(const MarkerRec& r) { return *reinterpret_cast<const int4*>(&r.i0); }

// Bucket of a record that the moving-marker map does not hold: -1 when it
// contributes nothing or when its base cell is the ordering's (bases[p]).
template <int D>
__device__ inline int moved_bucket(const int4& r, const Base* bases, int p, const Geo& g) {
  int b = rec_bucket<D>(r, g);
  if (b >= 0 && bases) {
    const Base o = bases[p];
    if (o.valid && o.i0 == r.x && o.j0 == r.y && (D == 2 || o.k0 == r.z)) b = -1;
  }
  return b;
}

// Marks in tflag the tiles of TileShape<D> (ntx x nty per z-layer) that the
// stencil nodes of a record with base cell r reach (mapped as axis_node).
template <int D>
__device__ inline void mark_tiles(const int4& r, const Geo& g, const Walls& w, int fold, unsigned char* tflag) {
  typedef TileShape<D> T;
  const int ntx = (g.nx + T::X - 1) / T::X, nty = (g.ny + T::Y - 1) / T::Y;
  int tx[4], ty[4], tz[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    const int ax = axis_node(r.x - 1 + q, g.nx, w.type[0][0] == 0, w.type[0][1] == 0, fold);
    const int ay = axis_node(r.y - 1 + q, g.ny, w.type[1][0] == 0, w.type[1][1] == 0, fold);
    const int az = (D == 3) ? axis_node(r.z - 1 + q, g.nz, w.type[2][0] == 0, w.type[2][1] == 0, fold) : 1;
    tx[q] = ax > 0 ? (ax - 1) / T::X : -1;
    ty[q] = ay > 0 ? (ay - 1) / T::Y : -1;
    tz[q] = az > 0 ? (az - 1) / T::Z : -1;
  }
#pragma unroll
  for (int c = 0; c < ((D == 3) ? 4 : 1); c++)
#pragma unroll
    for (int bb = 0; bb < 4; bb++)
#pragma unroll
      for (int a = 0; a < 4; a++)
        if (tx[a] >= 0 && ty[bb] >= 0 && tz[c] >= 0)
          tflag[(size_t)tx[a] + (size_t)ntx * ((size_t)ty[bb] + (size_t)nty * (size_t)tz[c])] = 1;
}

// The binning buffers of the records being spread (pass 1's outputs).
struct RecBin {
  int* cnt;              // per bucket
  int* bk;               // per record: bucket or -1
  int* rank;             // per record: arrival rank in the bucket
  unsigned char* tflag;  // when set: tiles reached by the binned records
  int fold;
};

// pass 1 for one record (slot s, marker p, base cell and validity r): its
// bucket and its arrival rank within the bucket; with ss.bases set only the
// records that left their ordered base cell are binned
template <int D>
__device__ inline void bin_record(const SlotSpan& ss, int s, int p, const int4& r, const Geo& g,
                                  const RecBin& bin) {
  const int b = moved_bucket<D>(r, ss.bases, p, g);
  const int e = ss.off[s] + p;
  bin.bk[e] = b;
  bin.rank[e] = (b >= 0) ? atomicAdd(bin.cnt + b, 1) : -1;
}

template <int D>
__global__ void __launch_bounds__(256) k_rec_count(SlotSpan ss, Geo g, RecBin bin) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ss.off[ss.nslots]) return;
  int s = 0;
  while (e >= ss.off[s + 1]) s++;
  const int p = e - ss.off[s];
  bin_record<D>(ss, s, p, rec_cell(ss.rec[s][p]), g, bin);
}

// the marked tiles as a list (any order), their marks cleared
// (thread 0 also copies the number of binned records, *nbinned, to the
// host-visible *report when given)
__global__ void __launch_bounds__(256) k_tile_list(int ntiles, unsigned char* tflag, int* list, int* count,
                                                   const int* nbinned = nullptr, int* report = nullptr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0 && report) *report = *nbinned;
  if (t >= ntiles || !tflag[t]) return;
  list[atomicAdd(count, 1)] = t;
  tflag[t] = 0;
}

// A binned record: (slot << 24 | marker) << 32 | (i0 + 2), so that sorting
// the entries of a bucket as integers orders them by (slot, marker).
__device__ inline unsigned long long entry_of(unsigned val, int i0) {
  return ((unsigned long long)val << 32) | (unsigned long long)(unsigned)(i0 + 2);
}
__device__ inline int entry_i0(unsigned long long e) { return (int)(unsigned)(e & 0xFFFFFFFFull) - 2; }
__device__ inline unsigned entry_val(unsigned long long e) { return (unsigned)(e >> 32); }

// pass 2: every contributing record into its bucket (arrival order within
// it); with tflag set, the tiles its stencil reaches are marked
template <int D>
__global__ void __launch_bounds__(256) k_rec_place(SlotSpan ss, Geo g, Walls w, int fold, const int* bk,
                                                  const int* rank, const int* bstart, unsigned long long* ent,
                                                  unsigned char* tflag) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ss.off[ss.nslots]) return;
  const int b = bk[e];
  if (b < 0) return;
  int s = 0;
  while (e >= ss.off[s + 1]) s++;
  const int p = e - ss.off[s];
  const int4 r = rec_cell(ss.rec[s][p]);
  ent[bstart[b] + rank[e]] = entry_of(((unsigned)s << REC_SLOT_SHIFT) | (unsigned)p, r.x);
  if (tflag) mark_tiles<D>(r, g, w, fold, tflag);
}

// pass 3: each bucket sorted: up to SB_SMALL entries in registers (all
// loads in flight together); larger buckets are listed for
// k_bucket_sort_large
constexpr int SB_SMALL = 8;
constexpr int SB_REG = SB_SMALL;
__global__ void __launch_bounds__(256) k_bucket_sort(int nb, const int* bstart, unsigned long long* ent,
                                                    int* large, int* nlarge) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int lo = bstart[b], n = bstart[b + 1] - lo;
  if (n <= 1) return;
  if (n <= SB_REG) {
    unsigned long long r[SB_REG];
#pragma unroll
    for (int i = 0; i < SB_REG; i++) r[i] = (i < n) ? ent[lo + i] : ~0ull;
    // odd-even transposition sort of the eight (padding sorts last)
#pragma unroll
    for (int pass = 0; pass < SB_REG; pass++) {
#pragma unroll
      for (int i = pass & 1; i + 1 < SB_REG; i += 2) {
        const unsigned long long x = r[i], y = r[i + 1];
        r[i] = x < y ? x : y;
        r[i + 1] = x < y ? y : x;
      }
    }
#pragma unroll
    for (int i = 0; i < SB_REG; i++)
      if (i < n) ent[lo + i] = r[i];
    return;
  }
  large[atomicAdd(nlarge, 1)] = b;
}

// large buckets: one block each, rank sort (entries are unique) into the
// second entry buffer, then copied back
__global__ void __launch_bounds__(256) k_bucket_sort_large(const int* large, const int* nlarge, const int* bstart,
                                                          unsigned long long* ent, unsigned long long* ent2) {
  const int nl = *nlarge;
  for (int i = blockIdx.x; i < nl; i += gridDim.x) {
    const int b = large[i];
    const int lo = bstart[b], n = bstart[b + 1] - lo;
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const unsigned long long v = ent[lo + e];
      int r = 0;
      for (int f = 0; f < n; f++) r += (ent[lo + f] < v) ? 1 : 0;
      ent2[lo + r] = v;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < n; e += blockDim.x) ent[lo + e] = ent2[lo + e];
    __syncthreads();
  }
}

struct TileSpreadArgs {
  Geo g;
  Walls w;
  const int* bstart;     // bucket starts, nb + 1 entries
  const unsigned long long* ent;  // binned records (entry_of), bucket by bucket
  SlotSpan ss;
  double* F4;            // node force (fx, fy, fz, unused)
  unsigned char* flag;   // 1 where some record contributes
  int fold;              // 1: periodic ghost images add into the interior (exchange_forces)
  int accumulate;        // 1: add to the forces already in F4 where flag is set
  int serial;            // 1: a periodic axis of fewer than 4 nodes (one record may reach
                         //    one node twice): lanes add one at a time
  const int* tiles;      // when set: the tiles to spread (tile list, *ntiles of them);
  const int* ntiles;     //   F4 written and flag set to tag only where some record contributes
  unsigned char tag;
};

// Base-cell values of one axis that a tile of ghost-inclusive nodes lo..hi
// needs: [lo - 2, hi + 1], then, when folding, those of the ghost-plane
// images not already listed (0 when the tile holds node n and the lo face is
// periodic: [-2, 1]; n + 1 when it holds node 1 and the hi face is periodic:
// [n - 1, n + 2]).  Written by the whole wavefront into out (at most
// hi - lo + 12 values); returns the count.
__device__ inline int axis_cells(int lo, int hi, int n, int tlo, int thi, int fold, int* out) {
  const int lane = threadIdx.x & 63;
  const int m = hi - lo + 4;
  if (lane < m) out[lane] = lo - 2 + lane;
  const bool elo = fold && hi >= n && tlo == 0, ehi = fold && lo <= 1 && thi == 0;
  int c = 0;
  bool ok = false;
  if (lane < 4) {
    c = -2 + lane;
    ok = elo;
  } else if (lane < 8) {
    c = n - 1 + (lane - 4);
    ok = ehi && !(elo && c >= -2 && c <= 1);
  }
  ok = ok && (c < lo - 2 || c > hi + 1);
  const unsigned long long bm = __ballot(ok);
  if (ok) out[m + __popcll(bm & ((1ull << lane) - 1ull))] = c;
  return m + __popcll(bm);
}


constexpr int SP_AXIS = 32;  // base-cell values per axis (tile extent + 3 + 8 at most)
constexpr int SP_NV = 16;    // doubles staged per record: wx[4], wy[4], wz[4], sc, fx, fy, fz
constexpr int SP_STAGE = 32; // records staged per round

// One wavefront (64 threads) per tile of TX x TY x TZ interior nodes.  The
// records that can reach the tile are those in the buckets of the base rows
// (j0, k0) within two cells of the tile (and of the folded ghost images) whose
// i0 window meets the tile's; they are enumerated per (base row, x-window)
// in a fixed order, filtered by i0, staged 64 at a time into LDS and walked
// in that order.  For each record lane (t, u, v) handles stencil node
// (i0 - 1 + t, j0 - 1 + u, k0 - 1 + v): the interior node it stands for (the
// node, or the node whose periodic ghost image it is) gets
// F * ((wx[t] wy[u]) wz[v]) * sc added into its LDS accumulator if it lies in
// the tile.
template <int D>
__global__ void __launch_bounds__(64) k_spread_tiles(TileSpreadArgs a) {
#pragma clang fp contract(off)
  constexpr int TX = TileShape<D>::X, TY = TileShape<D>::Y, TZ = TileShape<D>::Z;
  constexpr int PX = TileShape<D>::PX, PY = TileShape<D>::PY;
  constexpr int NT = TX * TY * TZ, NA = PX * PY * TZ;
  __shared__ double acc[3][NA];
  __shared__ double rwt[SP_STAGE][12];  // wx[4], wy[4], wz[4] of the staged records
  __shared__ double4 rsf[SP_STAGE];     // sc, fx, fy, fz
  __shared__ int4 rc[SP_STAGE];  // i0, j0, k0 of the staged records
  __shared__ int Yl[SP_AXIS], Zl[SP_AXIS];
  __shared__ int pst[64], poff[65], pwin[64];
  const Geo& g = a.g;
  const int lane = threadIdx.x;
  // the tile of this workgroup, or the listed tiles one after the other
  const int ntl = a.tiles ? *a.ntiles : 1;
  for (int it = a.tiles ? (int)blockIdx.x : 0; it < ntl; it += a.tiles ? (int)gridDim.x : 1) {
    int bx = blockIdx.x, by = blockIdx.y, bz = blockIdx.z;
    if (a.tiles) {
      const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY;
      const int tl = a.tiles[it];
      bx = tl % ntx;
      by = (tl / ntx) % nty;
      bz = tl / (ntx * nty);
    }
    const int X0 = bx * TX, Y0 = by * TY, Z0 = bz * TZ;
    const int xl = (g.nx - X0 < TX) ? g.nx - X0 : TX;
    const int yl = (g.ny - Y0 < TY) ? g.ny - Y0 : TY;
    const int zl = (D == 3) ? ((g.nz - Z0 < TZ) ? g.nz - Z0 : TZ) : 1;
    for (int i = lane; i < NA; i += 64) {
      acc[0][i] = 0.0;
      acc[1][i] = 0.0;
      acc[2][i] = 0.0;
    }
    const int nY = axis_cells(Y0 + 1, Y0 + yl, g.ny, a.w.type[1][0], a.w.type[1][1], a.fold, Yl);
    int nZ = 1;
    if (D == 3)
      nZ = axis_cells(Z0 + 1, Z0 + zl, g.nz, a.w.type[2][0], a.w.type[2][1], a.fold, Zl);
    else if (lane == 0)
      Zl[0] = 0;
    // x-windows of i0, in increasing order: that of the ghost image of node nx
    // (records over ghost 0: [-2, 1]), the tile's own, that of the ghost image
    // of node 1 (records over ghost nx + 1: [nx - 1, nx + 2]); overlapping or
    // touching windows merged so that no record is taken twice
    int wl0 = X0 - 1, wh0 = X0 + xl + 1, wl1 = 0, wh1 = -1, wl2 = 0, wh2 = -1, nw = 1;
    if (a.fold && a.w.type[0][0] == 0 && X0 + xl == g.nx) {
      if (wl0 <= 2) {
        wl0 = -2;
      } else {
        wl1 = wl0; wh1 = wh0; wl0 = -2; wh0 = 1; nw = 2;
      }
    }
    if (a.fold && a.w.type[0][1] == 0 && X0 == 0) {
      const int lo = g.nx - 1, hi = g.nx + 2;
      if (nw == 1) {
        if (lo <= wh0 + 1) wh0 = hi > wh0 ? hi : wh0;
        else { wl1 = lo; wh1 = hi; nw = 2; }
      } else {
        if (lo <= wh1 + 1) wh1 = hi > wh1 ? hi : wh1;
        else { wl2 = lo; wh2 = hi; nw = 3; }
      }
    }
    wh0 = wh0 > g.nx + 2 ? g.nx + 2 : wh0;
    wh1 = wh1 > g.nx + 2 ? g.nx + 2 : wh1;
    wh2 = wh2 > g.nx + 2 ? g.nx + 2 : wh2;
    __syncthreads();
    const int nxb = spread_nxb(g);
    // this lane's stencil offsets and the per-axis face kinds
    const int t = lane & 3, u = (lane >> 2) & 3, v = lane >> 4;
    const bool lane_on = (D == 3) || v == 0;
    const bool pxl = a.w.type[0][0] == 0, pxh = a.w.type[0][1] == 0;
    const bool pyl = a.w.type[1][0] == 0, pyh = a.w.type[1][1] == 0;
    const bool pzl = a.w.type[2][0] == 0, pzh = a.w.type[2][1] == 0;
    const int npair = nY * nZ * nw;  // pair = (zi * nY + yi) * nw + w
    for (int p0 = 0; p0 < npair; p0 += 64) {
      // this lane's (base row, window): its range of binned entries
      const int p = p0 + lane;
      int e0 = 0, cnt = 0, w = 0;
      if (p < npair) {
        w = p % nw;
        const int bi = p / nw;
        const int yi = bi % nY, zi = bi / nY;
        const int row = (Yl[yi] + 2) + g.nry * ((D == 3) ? Zl[zi] + 2 : 0);
        const int wlo = (w == 0) ? wl0 : (w == 1) ? wl1 : wl2, whi = (w == 0) ? wh0 : (w == 1) ? wh1 : wh2;
        const int b0 = row * nxb + (wlo + 2) / SB_X, b1 = row * nxb + (whi + 2) / SB_X;
        e0 = a.bstart[b0];
        cnt = a.bstart[b1 + 1] - e0;
      }
      // inclusive scan of the counts over the wavefront
      int incl = cnt;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
      }
      const int total = __shfl(incl, 63);
      pst[lane] = e0;
      poff[lane] = incl - cnt;
      pwin[lane] = w;
      if (lane == 0) poff[64] = total;
      __syncthreads();
      const int npl = (npair - p0 < 64) ? npair - p0 : 64;
      for (int s0 = 0; s0 < total; s0 += 64) {
        // one entry per lane: its pair by binary search over the offsets,
        // kept if its i0 lies in the pair's window
        const int gq = s0 + lane;
        bool keep = false;
        unsigned long long en = 0;
        if (gq < total) {
          int lo = 0, hi = npl - 1;
          while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (poff[mid] <= gq) lo = mid; else hi = mid - 1;
          }
          en = a.ent[pst[lo] + (gq - poff[lo])];
          const int i0 = entry_i0(en);
          const int ww = pwin[lo];
          const int wlo = (ww == 0) ? wl0 : (ww == 1) ? wl1 : wl2, whi = (ww == 0) ? wh0 : (ww == 1) ? wh1 : wh2;
          keep = i0 >= wlo && i0 <= whi;
        }
        const unsigned long long km = __ballot(keep);
        const int nkeep = __popcll(km);
        const int q = keep ? __popcll(km & ((1ull << lane) - 1ull)) : 0;
        // the kept entries in rounds of SP_STAGE staged records
        for (int r0 = 0; r0 < nkeep; r0 += SP_STAGE) {
          const int nst = (nkeep - r0 < SP_STAGE) ? nkeep - r0 : SP_STAGE;
          if (keep && q >= r0 && q < r0 + SP_STAGE) {
            const unsigned vv = entry_val(en);
            const MarkerRec* r = a.ss.rec[vv >> REC_SLOT_SHIFT] + (vv & ((1u << REC_SLOT_SHIFT) - 1u));
            // the record's first 16 doubles and its base cell (records are 144
            // bytes, so 16-byte aligned)
            const double2* rp = reinterpret_cast<const double2*>(r);
            double2 rv[SP_NV / 2];
#pragma unroll
            for (int m = 0; m < SP_NV / 2; m++) rv[m] = rp[m];
            int4 ci = *reinterpret_cast<const int4*>(&r->i0);
#pragma unroll
            for (int m = 0; m < 6; m++) {
              rwt[q - r0][2 * m] = rv[m].x;
              rwt[q - r0][2 * m + 1] = rv[m].y;
            }
            rsf[q - r0] = double4{rv[6].x, rv[6].y, rv[7].x, rv[7].y};
            // w: the whole stencil lies in the interior (no ghost node)
            ci.w = (ci.x >= 2 && ci.x + 2 <= g.nx && ci.y >= 2 && ci.y + 2 <= g.ny &&
                    (D == 2 || (ci.z >= 2 && ci.z + 2 <= g.nz))) ? 1 : 0;
            rc[q - r0] = ci;
          }
          __syncthreads();
          for (int e = 0; e < nst; e++) {
            const int4 ci = rc[e];
            // the interior node this lane's stencil node stands for, in the tile
            int lx, ly, lz;
            bool in;
            if (ci.w) {
              lx = ci.x - 2 + t - X0;
              ly = ci.y - 2 + u - Y0;
              lz = (D == 3) ? ci.z - 2 + v - Z0 : 0;
              in = lane_on;
            } else {
              const int ax = axis_node(ci.x - 1 + t, g.nx, pxl, pxh, a.fold);
              const int ay = axis_node(ci.y - 1 + u, g.ny, pyl, pyh, a.fold);
              const int az = (D == 3) ? axis_node(ci.z - 1 + v, g.nz, pzl, pzh, a.fold) : 1;
              lx = ax - 1 - X0;
              ly = ay - 1 - Y0;
              lz = az - 1 - Z0;
              in = lane_on && ax > 0 && ay > 0 && az > 0;
            }
            in = in && (unsigned)lx < (unsigned)TX && (unsigned)ly < (unsigned)TY && (unsigned)lz < (unsigned)TZ;
            if (in) {
              const int idx = lx + PX * (ly + PY * lz);
              const double d = (D == 3) ? (rwt[e][t] * rwt[e][4 + u]) * rwt[e][8 + v] : rwt[e][t] * rwt[e][4 + u];
              const double4 sf = rsf[e];  // sc, fx, fy, fz
              const double wgt = d * sf.x;
              const double fx = sf.y * wgt, fy = sf.z * wgt, fz = sf.w * wgt;
              if (!a.serial) {
                acc[0][idx] += fx;
                acc[1][idx] += fy;
                acc[2][idx] += fz;
              } else {
                for (int L = 0; L < 64; L++) {
                  if (lane == L) {
                    acc[0][idx] += fx;
                    acc[1][idx] += fy;
                    acc[2][idx] += fz;
                  }
                  __builtin_amdgcn_wave_barrier();
                }
              }
            }
          }
          __syncthreads();  // the next round overwrites the staged records
        }
      }
    }
    __syncthreads();
    for (int i = lane; i < NT; i += 64) {
      const int lx = i % TX, ly = (i / TX) % TY, lz = i / (TX * TY);
      if (lx >= xl || ly >= yl || lz >= zl) continue;
      const size_t node = (size_t)(X0 + lx) + (size_t)g.nx * ((size_t)(Y0 + ly) + (size_t)g.ny * (size_t)(Z0 + lz));
      const int ai = lx + PX * (ly + PY * lz);
      const double ax = acc[0][ai], ay = acc[1][ai], az = acc[2][ai];
      // a node is flagged where its force is nonzero: the sums start at +0.0,
      // so an unflagged node and a flagged zero force give the LB step the same
      // +0.0 (the reference's cleared grid force plus what was spread)
      const bool h = ax != 0.0 || ay != 0.0 || az != 0.0;
      if (a.tiles) {
        if (h) {
          *reinterpret_cast<double4*>(a.F4 + 4 * node) = double4{ax, ay, az, 0.0};
          a.flag[node] = a.tag;
        }
      } else if (a.accumulate) {
        if (h) {
          double4* F = reinterpret_cast<double4*>(a.F4 + 4 * node);
          if (a.flag[node]) {
            const double4 o = *F;
            *F = double4{o.x + ax, o.y + ay, o.z + az, 0.0};
          } else {
            *F = double4{ax, ay, az, 0.0};
          }
          a.flag[node] = 1;
        }
      } else {
        a.flag[node] = h ? 1 : 0;
        if (h) *reinterpret_cast<double4*>(a.F4 + 4 * node) = double4{ax, ay, az, 0.0};
      }
    }
    __syncthreads();  // the next tile clears the accumulators
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
