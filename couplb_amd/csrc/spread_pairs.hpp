// spread_pairs.hpp -- force spreading through a node-sorted map of (marker,
// stencil node) pairs: for markers that keep their positions from one LB
// interval to the next, and for moving markers that keep their base cells.
//
// The weights of IBM::spread (couplb_ibm.h:92-133) depend on the marker
// positions only, so for fixed markers the map from marker reactions to node
// forces is a constant sparse matrix.  It is built once per marker ordering:
// every (marker, stencil node) pair that lands on a fluid node is keyed by
// the interior node it adds into (periodic ghost images folded as
// exchange_forces does, couplb_streaming.h:217-225), the pairs are stably
// radix-sorted by node, and each keeps its weight delta*sc.  Per LB step the
// first pair of every node's run sums the run in sorted order (ascending
// marker, then stencil slot): a deterministic segmented reduction that reads
// 16 bytes per pair plus the marker forces.
//
// Moving markers: the pairs depend on the base cells only, so the map of the
// marker ordering's base cells (k_pairs_base) holds for every record whose
// marker has kept its base cell; per LB step k_pair_con sums each (marker,
// stencil node)'s force over the interval's records and the LB step sums a
// node's run through the sorted pair values.  Records of markers that left
// their base cell are spread by spread_rows.hpp.
#pragma once
#include "kernels.hpp"
#include "spread_rows.hpp"

namespace couplb {

// one thread per (marker p, stencil slot s); slot s = t + 4u + 16v is
// ghost-inclusive node (i0-1+t, j0-1+u, k0-1+v) as in IBM::spread
template <int D>
__global__ void __launch_bounds__(256) k_pairs(int n, Geo g, Walls w, const MarkerRec* rec,
                                               unsigned sentinel, unsigned* key, int* val) {
  constexpr int S = stencil_size<D>();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * S) return;
  const int p = (int)(e / S), s = (int)(e % S);
  const MarkerRec& r = rec[p];
  unsigned k = sentinel;
  if (r.valid) {
    const int a = r.i0 - 1 + (s & 3), b = r.j0 - 1 + ((s >> 2) & 3);
    const int c = (D == 3) ? r.k0 - 1 + (s >> 4) : 0;
    const bool in = a >= 0 && a <= g.nx + 1 && b >= 0 && b <= g.ny + 1 &&
                    (D == 2 || (c >= 0 && c <= g.nz + 1));
    int wd, wsd;
    if (in && node_type(w, g, a, b, c, wd, wsd) == 0)
      k = (unsigned)wrap_ghost(a, g.nx) +
          (unsigned)g.nx * ((unsigned)wrap_ghost(b, g.ny) +
                            (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0));
  }
  key[e] = k;
  val[e] = (int)e;
}

// The moving-marker map: the same pairs from the base cells of the marker
// ordering (Base, ibm.hpp) instead of one slot's records, so that it holds
// for every record whose base cell is unchanged since the ordering.
template <int D>
__global__ void __launch_bounds__(256) k_pairs_base(int n, Geo g, Walls w, const Base* bases,
                                                    unsigned sentinel, unsigned* key, int* val) {
  constexpr int S = stencil_size<D>();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)n * S) return;
  const int p = (int)(e / S), s = (int)(e % S);
  const Base b = bases[p];
  unsigned k = sentinel;
  if (b.valid) {
    const int a = b.i0 - 1 + (s & 3), bb = b.j0 - 1 + ((s >> 2) & 3);
    const int c = (D == 3) ? b.k0 - 1 + (s >> 4) : 0;
    const bool in = a >= 0 && a <= g.nx + 1 && bb >= 0 && bb <= g.ny + 1 &&
                    (D == 2 || (c >= 0 && c <= g.nz + 1));
    int wd, wsd;
    if (in && node_type(w, g, a, bb, c, wd, wsd) == 0)
      k = (unsigned)wrap_ghost(a, g.nx) +
          (unsigned)g.nx * ((unsigned)wrap_ghost(bb, g.ny) +
                            (unsigned)g.ny * (unsigned)((D == 3) ? wrap_ghost(c, g.nz) : 0));
  }
  key[e] = k;
  val[e] = (int)e;
}

// after the sort of the moving-marker map: the node flags (the caller has
// cleared them)
__global__ void __launch_bounds__(256) k_map_flags(long ne, const unsigned* skey, unsigned sentinel,
                                                   unsigned char* flag) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const unsigned k = skey[e];
  if (k != sentinel && (e == 0 || skey[e - 1] != k)) flag[k] = 1;
}

// Force of every (marker p, stencil slot s) of the moving-marker map for this
// LB step, in marker order (con: four doubles per entry p * S + s, the last
// unused; the LB step reads a node's entries through the sorted pair
// values), summing over the pending record slots in sub-step order the
// records of p whose base cell is the map's: F * (((wx wy) wz) sc), as
// couplb_ibm.h:125-130 adds it.  One thread per (marker, stencil plane v) in
// 3D (per marker in 2D) computes the plane's 16 entries; the other records of
// p (it left its ordered base cell) are spread separately, and thread v also
// bins the marker's records v, v + 4, ... for that (spread_rows.hpp
// bin_record; ss.bases is the ordering's base cells).
template <int D>
__global__ void __launch_bounds__(256) k_pair_con(int n, SlotSpan ss, double* con, Geo g, RecBin bin) {
#pragma clang fp contract(off)
  constexpr int LPM = (D == 3) ? 4 : 1;  // threads per marker
  constexpr int S = stencil_size<D>();
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * LPM) return;
  const int p = tid / LPM, v = tid % LPM;
  const Base b = ss.bases[p];
  double c[16][3];
#pragma unroll
  for (int q = 0; q < 16; q++) c[q][0] = c[q][1] = c[q][2] = 0.0;
  for (int k = 0; k < ss.nslots; k++) {
    const MarkerRec& r = ss.rec[k][p];
    const int4 rb = *reinterpret_cast<const int4*>(&r.i0);
    if (k % LPM == v) bin_record<D>(ss, k, p, rb, g, bin);
    if (!rb.w || rb.x != b.i0 || rb.y != b.j0 || (D == 3 && rb.z != b.k0)) continue;
    double wx[4], wy[4];
#pragma unroll
    for (int t = 0; t < 4; t++) { wx[t] = r.wx[t]; wy[t] = r.wy[t]; }
    const double wz = (D == 3) ? r.wz[v] : 1.0, sc = r.sc, fx = r.fx, fy = r.fy, fz = r.fz;
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const double d = (D == 3) ? (wx[t] * wy[u]) * wz : wx[t] * wy[u];
        const double wgt = d * sc;
        c[4 * u + t][0] += fx * wgt;
        c[4 * u + t][1] += fy * wgt;
        c[4 * u + t][2] += fz * wgt;
      }
  }
  double4* o = reinterpret_cast<double4*>(con) + (size_t)p * S + 16 * v;
#pragma unroll
  for (int q = 0; q < 16; q++) o[q] = double4{c[q][0], c[q][1], c[q][2], 0.0};
}

// after the sort: marker index and weight of every pair, and the node flags
// (the caller has cleared them)
template <int D>
__global__ void __launch_bounds__(256) k_pair_weights(long ne, const unsigned* skey, const int* sval,
                                                      unsigned sentinel, const MarkerRec* rec,
                                                      int* pp, double* pw, unsigned char* flag) {
#pragma clang fp contract(off)
  constexpr int S = stencil_size<D>();
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const unsigned k = skey[e];
  if (k == sentinel) return;
  const int v = sval[e];
  const int p = v / S, s = v % S;
  const MarkerRec& r = rec[p];
  const double d = (D == 3) ? (r.wx[s & 3] * r.wy[(s >> 2) & 3]) * r.wz[s >> 4]
                            : r.wx[s & 3] * r.wy[(s >> 2) & 3];
  pp[e] = p;
  pw[e] = d * r.sc;  // w = dcache[i] * sc, couplb_ibm.h:126
  if (e == 0 || skey[e - 1] != k) flag[k] = 1;
}

// first pair of every node's run
__global__ void __launch_bounds__(256) k_pair_heads(long ne, const unsigned* skey, unsigned sentinel,
                                                    unsigned char* hflag) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const unsigned k = skey[e];
  hflag[e] = (k != sentinel && (e == 0 || skey[e - 1] != k)) ? 1 : 0;
}

// node and length of each run (heads in increasing pair order), for the
// *nh runs the selection counted (launched over a bound on their number);
// the per-node run length plen[node] for the LB step (StepArgs::plen, which
// the caller has cleared), and over[0] = 1 where a run is too long for it
__global__ void __launch_bounds__(256) k_head_info(long bound, const int* nh, const int* heads,
                                                   const unsigned* skey, long ne, int* hlen,
                                                   unsigned* hnode, unsigned short* plen, int* over) {
  const long h = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= bound || h >= (long)*nh) return;
  const long s = heads[h];
  const unsigned k = skey[s];
  long f = s + 1;
  while (f < ne && skey[f] == k) f++;
  hlen[h] = (int)(f - s);
  hnode[h] = k;
  if (f - s > 0xffff) over[0] = 1;
  plen[k] = (unsigned short)(f - s > 0xffff ? 0xffff : f - s);
}

// first pair of each 64-node segment of an x-row (segment sg = row * nsx +
// x / 64, nsx = ceil(nx / 64)): the first sorted pair whose node is at or
// past the segment's first node; with the run lengths plen it locates every
// node's run (kernels.hpp pair_run)
__global__ void __launch_bounds__(256) k_seg_start(long nseg, int nsx, int nx, const unsigned* skey, long ne,
                                                   int* pseg) {
  const long sg = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (sg >= nseg) return;
  const unsigned n0 = (unsigned)(sg / nsx) * (unsigned)nx + 64u * (unsigned)(sg % nsx);
  long lo = 0, hi = ne;  // lower bound of n0 in skey[0..ne)
  while (lo < hi) {
    const long m = (lo + hi) >> 1;
    if (skey[m] < n0) lo = m + 1;
    else hi = m;
  }
  pseg[sg] = (int)lo;
}

// p[0..n) = v, 16 bytes per thread (p 16-byte aligned, as device
// allocations are); launch with fill_u8_blocks(n) blocks of 256 threads
__global__ void k_fill_u8(unsigned char* p, size_t n, unsigned char v) {
  typedef unsigned int V4 __attribute__((ext_vector_type(4)));
  const size_t i = ((size_t)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i + 16 <= n) {
    const unsigned w = (unsigned)v * 0x01010101u;
    *reinterpret_cast<V4*>(p + i) = V4{w, w, w, w};
  } else {
    for (size_t k = i; k < n; k++) p[k] = v;
  }
}
inline unsigned fill_u8_blocks(size_t n) { return (unsigned)((n + 16 * 256 - 1) / (16 * 256)); }

}  // namespace couplb
