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

template <int D> __host__ __device__ constexpr int stencil_size() { return D == 3 ? 64 : 16; }

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

// F_node = sum over the node's run of F_marker * w (couplb_ibm.h:125-130)
__global__ void __launch_bounds__(256) k_pair_spread(long ne, const unsigned* skey, const int* pp,
                                                     const double* pw, unsigned sentinel,
                                                     const MarkerRec* rec, double* Fx, double* Fy,
                                                     double* Fz) {
#pragma clang fp contract(off)
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const unsigned k = skey[e];
  if (k == sentinel || (e > 0 && skey[e - 1] == k)) return;
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (long f = e; f < ne && skey[f] == k; f++) {
    const MarkerRec& r = rec[pp[f]];
    const double wgt = pw[f];
    ax += r.fx * wgt;
    ay += r.fy * wgt;
    az += r.fz * wgt;
  }
  Fx[k] = ax;
  Fy[k] = ay;
  Fz[k] = az;
}

__global__ void k_fill_u8(unsigned char* p, size_t n, unsigned char v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace couplb
