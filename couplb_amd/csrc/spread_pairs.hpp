// spread_pairs.hpp -- force spreading for markers that keep their positions
// from one LB interval to the next.
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
#pragma once
#include "kernels.hpp"

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

// node and length of each run (heads in increasing pair order)
__global__ void __launch_bounds__(256) k_head_info(int nh, const int* heads, const unsigned* skey,
                                                   long ne, int* hlen, unsigned* hnode) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const long s = heads[h];
  const unsigned k = skey[s];
  long f = s + 1;
  while (f < ne && skey[f] == k) f++;
  hlen[h] = (int)(f - s);
  hnode[h] = k;
}

// run of every flagged node: prun[node] = (first pair, pair count), read by
// the LB step (StepArgs::pair_mode)
__global__ void k_run_index(int nh, const int* heads, const int* hlen, const unsigned* hnode, int2* prun) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  prun[hnode[h]] = int2{heads[h], hlen[h]};
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
