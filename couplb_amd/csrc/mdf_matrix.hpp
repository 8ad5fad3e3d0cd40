// mdf_matrix.hpp -- the linear part of the multi-direct-forcing operator as
// an explicit sparse marker-by-marker matrix.
//
//   (A p)_k = I0(S(p)/(2 rho))_k = g_k sum_x pw_kx / (2 rho_x) sum_j pw_jx p_j
//
// with pw_jx the spreading weight delta*sc of marker j at node x (the sorted
// pairs of spread_pairs.hpp), and g_k = (fluid weight)/(all counted weight)
// of marker k, which turns pw into the interpolation weight delta/ws of
// IBM::interpolate (couplb_ibm.h:57-90).  A_kj is non-zero where the two
// stencils share a fluid node.  The structure (which (k, j) and which shared
// nodes) depends on the marker cells only and is built once per ordering:
// every node run emits its len^2 (k, j) triples, a stable radix sort groups
// them by (k, j), and each matrix entry keeps its list of triples.  The
// values depend on the density and are summed per solve from those lists
// in their fixed order.  A CG pass is then one sparse matrix-vector product
// instead of a spreading and an interpolation over the grid.
#pragma once
#include "mdf.hpp"

namespace couplb {

// triples per node run (len^2), and a zero after the last run for the scan
__global__ void __launch_bounds__(256) k_mm_count(int nh, const int* hlen, unsigned long long* cnt) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h > nh) return;
  cnt[h] = h < nh ? (unsigned long long)hlen[h] * (unsigned long long)hlen[h] : 0ull;
}

// triples of node run h: key (k, j) = k*n + j, value (pair of k, pair of j)
__global__ void __launch_bounds__(256) k_mm_emit(int nh, const int* heads, const int* hlen,
                                                 const int* pp, unsigned long long n,
                                                 const unsigned long long* off,
                                                 unsigned long long* key, unsigned long long* val) {
  const int h = blockIdx.x * blockDim.x + threadIdx.x;
  if (h >= nh) return;
  const int s = heads[h], len = hlen[h];
  unsigned long long o = off[h];
  for (int u = 0; u < len; u++) {
    const unsigned long long k = (unsigned long long)pp[s + u];
    for (int v = 0; v < len; v++, o++) {
      key[o] = k * n + (unsigned long long)pp[s + v];
      val[o] = ((unsigned long long)(s + u) << 32) | (unsigned long long)(unsigned)(s + v);
    }
  }
}

__global__ void __launch_bounds__(256) k_mm_heads(long long T, const unsigned long long* skey,
                                                  unsigned char* flag) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= T) return;
  flag[t] = (t == 0 || skey[t - 1] != skey[t]) ? 1 : 0;
}

// per entry: column, first triple, triple count, row
__global__ void __launch_bounds__(256) k_mm_entries(int nnz, const int* ent,
                                                    const unsigned long long* skey, long long T,
                                                    unsigned long long n, int* col, int* erow,
                                                    int* elen) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const unsigned long long key = skey[ent[e]];
  col[e] = (int)(key % n);
  erow[e] = (int)(key / n);
  elen[e] = (int)((e + 1 < nnz ? (long long)ent[e + 1] : T) - (long long)ent[e]);
}

// rowptr[k] = first entry of row k (rows ascending), k = 0..n
__global__ void __launch_bounds__(256) k_mm_rowptr(int n, int nnz, const int* erow, int* rowptr) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > n) return;
  int lo = 0, hi = nnz;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (erow[mid] < k) lo = mid + 1; else hi = mid;
  }
  rowptr[k] = lo;
}

// values for the current density: A_kj = g_k sum_t pw[a_t] pw[b_t] / (2 rho(node of a_t))
__global__ void __launch_bounds__(256) k_mm_values(int nnz, const int* ent, const int* elen,
                                                   const int* erow,
                                                   const unsigned long long* sval,
                                                   const double* pw, const unsigned* pkey_s,
                                                   const double* u4, const double* gam,
                                                   double* aval) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  const int s = ent[e];
  double acc = 0.0;
  for (int t = 0; t < elen[e]; t++) {
    const unsigned long long v = sval[s + t];
    const unsigned a = (unsigned)(v >> 32), b = (unsigned)(v & 0xffffffffull);
    const double rho = u4[4 * (size_t)pkey_s[a] + 3];
    acc += pw[a] * pw[b] * (0.5 / rho);
  }
  aval[e] = gam[erow[e]] * acc;
}

// y = A x, eight lanes per row, summed in a fixed tree.
//   mode 0: Ap = y, block partial of x.y
//   mode 1: R = B - y, block partials of R.R and max |R|
constexpr int MM_LANES = 8;
__global__ void __launch_bounds__(MDF_THREADS) k_mm_spmv(int n, const int* rowptr, const int* col,
                                                         const double* aval, const double* x,
                                                         const double* B, double* y, int mode,
                                                         double* part_a, double* part_b,
                                                         const MdfScalars* S) {
  __shared__ double sh[MDF_THREADS / 64];
  if (S->done) return;
  const int k = (blockIdx.x * blockDim.x + threadIdx.x) / MM_LANES;
  const int l = threadIdx.x % MM_LANES;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  if (k < n) {
    // lane l takes entries l, l + 8, ... of the row, four loads in flight
    const int e1 = rowptr[k + 1];
    for (int e = rowptr[k] + l; e < e1; e += 4 * MM_LANES) {
      double v[4];
      int cc[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int ei = e + i * MM_LANES;
        const bool in = ei < e1;
        v[i] = in ? aval[ei] : 0.0;
        cc[i] = in ? col[ei] : k;
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        s0 += v[i] * x[3 * (size_t)cc[i] + 0];
        s1 += v[i] * x[3 * (size_t)cc[i] + 1];
        s2 += v[i] * x[3 * (size_t)cc[i] + 2];
      }
    }
  }
#pragma unroll
  for (int o = MM_LANES / 2; o > 0; o >>= 1) {
    s0 += __shfl_down(s0, o, MM_LANES);
    s1 += __shfl_down(s1, o, MM_LANES);
    s2 += __shfl_down(s2, o, MM_LANES);
  }
  double a = 0.0, b = 0.0;
  if (k < n && l == 0) {
    if (mode == 0) {
      y[3 * (size_t)k + 0] = s0;
      y[3 * (size_t)k + 1] = s1;
      y[3 * (size_t)k + 2] = s2;
      a = x[3 * (size_t)k + 0] * s0 + x[3 * (size_t)k + 1] * s1 + x[3 * (size_t)k + 2] * s2;
    } else {
      const double r0 = B[3 * (size_t)k + 0] - s0, r1 = B[3 * (size_t)k + 1] - s1,
                   r2 = B[3 * (size_t)k + 2] - s2;
      y[3 * (size_t)k + 0] = r0;
      y[3 * (size_t)k + 1] = r1;
      y[3 * (size_t)k + 2] = r2;
      a = r0 * r0 + r1 * r1 + r2 * r2;
      b = sqrt(a);
    }
  }
  a = block_reduce(a, 0, sh);
  b = block_reduce(b, 1, sh);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = a;
    part_b[blockIdx.x] = b;
  }
}

}  // namespace couplb
