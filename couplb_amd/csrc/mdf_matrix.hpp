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
//
// Block-Jacobi preconditioner (COUPLB_MDF_PCG): the markers are grouped in
// blocks of BJ along a Morton curve of their base cells, so that markers
// whose stencils overlap mostly share a block.  With about one marker per
// dx and a 4-dx stencil the directions plain CG resolves slowest are forces
// alternating between neighbouring markers; they live inside the blocks.
// Per solve each block's diagonal part of the symmetric factor
// Ah = G^-1 A (Ah_kj = sum_x pw_kx pw_jx / (2 rho_x), G = diag(g_k)) is
// inverted in LDS, and preconditioned CG applies z = blockdiag(Ah_bb)^-1 r
// in the pass that updates r.  The preconditioner changes the path of the
// iteration, not the system it solves.
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

// per pair a of the sorted map, its weight over twice the density of its
// node for the current state: q_a = pw_a (0.5 / rho(node of a))
__global__ void __launch_bounds__(256) k_mm_q(long ne, const double* pw, const unsigned* pkey_s,
                                              unsigned sentinel, const double* u4, double* q) {
  const long a = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ne) return;
  const unsigned node = pkey_s[a];
  q[a] = node != sentinel ? pw[a] * (0.5 / u4[4 * (size_t)node + 3]) : 0.0;
}

// values for the current density: A_kj = g_k sum_t q[a_t] pw[b_t].
// MV_LANES lanes per entry: lane l sums the entry's triples l, l + MV_LANES,
// ... in order (two loaded ahead), and the lanes' sums are added in a fixed
// tree, so the entries do not depend on scheduling.
constexpr int MV_LANES = 8;
__global__ void __launch_bounds__(256) k_mm_values(int nnz, const int* ent, const int* elen,
                                                   const int* erow,
                                                   const unsigned long long* sval,
                                                   const double* pw, const double* q,
                                                   const double* gam, double* aval) {
  const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int e = (int)(tid / MV_LANES), l = (int)(tid % MV_LANES);
  double acc = 0.0;
  if (e < nnz) {
    const int s = ent[e], len = elen[e];
    for (int t = l; t < len; t += 2 * MV_LANES) {
      const bool two = t + MV_LANES < len;
      const unsigned long long v0 = sval[s + t];
      const unsigned long long v1 = two ? sval[s + t + MV_LANES] : 0ull;
      const double x0 = q[(unsigned)(v0 >> 32)], y0 = pw[(unsigned)(v0 & 0xffffffffull)];
      const double x1 = two ? q[(unsigned)(v1 >> 32)] : 0.0;
      const double y1 = two ? pw[(unsigned)(v1 & 0xffffffffull)] : 0.0;
      acc += x0 * y0;
      if (two) acc += x1 * y1;
    }
  }
  // the lanes' sums in a fixed tree (the lanes of an entry are adjacent in a wave)
#pragma unroll
  for (int o = 1; o < MV_LANES; o <<= 1) acc += __shfl_down(acc, o, MV_LANES);
  if (e < nnz && l == 0) aval[e] = gam[erow[e]] * acc;
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

// ---- block-Jacobi preconditioner ----
constexpr int BJ = 32;  // markers per block

// 21 bits of v spread to every third bit
__device__ inline unsigned long long bj_spread(unsigned v) {
  unsigned long long m = 0;
  for (int b = 0; b < 21; b++) m |= (unsigned long long)((v >> b) & 1u) << (3 * b);
  return m;
}

// Morton key of each sorted marker's base cell (offset by 2: a base cell
// can sit on the ghost layer); markers without a stencil go last
__global__ void __launch_bounds__(256) k_bj_keys(int n, const MarkerRec* rec, unsigned long long* key,
                                                 int* val) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const MarkerRec& r = rec[k];
  key[k] = r.valid ? bj_spread((unsigned)(r.i0 + 2)) | (bj_spread((unsigned)(r.j0 + 2)) << 1) |
                         (bj_spread((unsigned)(r.k0 + 2)) << 2)
                   : (~0ull >> 1);
  val[k] = k;
}

// block order: bperm[t] = marker at position t (-1 past n), bpos = inverse
__global__ void __launch_bounds__(256) k_bj_perm(int n, int npad, const int* sval, int* bperm, int* bpos) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npad) return;
  const int k = t < n ? sval[t] : -1;
  bperm[t] = k;
  if (k >= 0) bpos[k] = t;
}

// One 64-lane block per marker block: gather the block's entries of Ah from
// the explicit operator (rows with no fluid weight decoupled: identity),
// invert by Gauss-Jordan elimination in LDS (Ah_bb is symmetric positive
// definite: no pivoting), write the BJ x BJ inverse column-major (element
// (r, q) at q * BJ + r: the lanes of a block read one column together).  A pivot
// that is not positive (a singular block) leaves the block's diagonal
// inverse instead.
__global__ void __launch_bounds__(64) k_bj_invert(const int* bperm, const int* bpos, const int* rowptr,
                                                  const int* col, const double* aval, const double* gam,
                                                  double* binv) {
  __shared__ double M[BJ][BJ + 1];
  __shared__ double dg[BJ];
  __shared__ int dz[BJ];
  const int b = blockIdx.x, l = threadIdx.x;
  for (int e = l; e < BJ * BJ; e += 64) M[e / BJ][e % BJ] = 0.0;
  __syncthreads();
  {
    // two lanes per row, each taking every other entry of the row, four
    // entries in flight (every entry has its own cell: the order of the
    // stores does not matter)
    const int r = l >> 1;
    const int k = bperm[b * BJ + r];
    const double gk = (k >= 0) ? gam[k] : 0.0;
    if (k >= 0 && gk > 0.0) {
      const int e1 = rowptr[k + 1];
      int e = rowptr[k] + (l & 1);
      for (; e + 6 < e1; e += 8) {
        int cj[4];
        double av[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          cj[u] = col[e + 2 * u];
          av[u] = aval[e + 2 * u];
        }
        int pj[4];
#pragma unroll
        for (int u = 0; u < 4; u++) pj[u] = bpos[cj[u]] - b * BJ;
#pragma unroll
        for (int u = 0; u < 4; u++)
          if (pj[u] >= 0 && pj[u] < BJ) M[r][pj[u]] = av[u] / gk;
      }
      for (; e < e1; e += 2) {
        const int pj = bpos[col[e]] - b * BJ;
        if (pj >= 0 && pj < BJ) M[r][pj] = aval[e] / gk;
      }
    }
  }
  __syncthreads();
  if (l < BJ) {
    const int k = bperm[b * BJ + l];
    const double d = (k >= 0 && gam[k] > 0.0) ? M[l][l] : 0.0;
    dz[l] = !(d > 0.0);
  }
  __syncthreads();
  for (int e = l; e < BJ * BJ; e += 64) {
    const int r = e / BJ, q = e % BJ;
    if (dz[r] || dz[q]) M[r][q] = (r == q) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (l < BJ) dg[l] = M[l][l];
  __syncthreads();
  const int row = l >> 1, c0 = (l & 1) * (BJ / 2);
  bool ok = true;
  for (int p = 0; p < BJ; p++) {
    const double piv = M[p][p];
    const double f = M[row][p];
    __syncthreads();
    if (!(piv > 1e-12 * dg[p]) || !(piv < 1e300)) {  // the same value in every lane
      ok = false;
      break;
    }
    // the pivot row, one column per lane
    if (l < BJ) M[p][l] = (l == p ? 1.0 : M[p][l]) / piv;
    __syncthreads();
    if (row != p) {  // half a row per lane, its values and the pivot row's in registers
      double pr[BJ / 2], mr[BJ / 2];
#pragma unroll
      for (int u = 0; u < BJ / 2; u++) {
        pr[u] = M[p][c0 + u];
        mr[u] = M[row][c0 + u];
      }
#pragma unroll
      for (int u = 0; u < BJ / 2; u++) M[row][c0 + u] = (c0 + u == p ? 0.0 : mr[u]) - f * pr[u];
    }
    __syncthreads();
  }
  double* o = binv + (size_t)b * BJ * BJ;
  for (int e = l; e < BJ * BJ; e += 64) {
    const int q = e / BJ, r = e % BJ;
    o[e] = ok ? M[r][q] : (r == q ? 1.0 / dg[r] : 0.0);
  }
}

// Preconditioned CG, markers in block order (MDF_THREADS / BJ blocks per
// workgroup), z = M^-1 r with the block's r staged in LDS:
//   UPDATE  alpha = rz / p.Ap; F += alpha p, r -= alpha Ap; then z (as
//           k_cg_update; p.Ap <= 0 leaves everything for k_cg_next)
//   restart z of the true residual r and p = z
// Block partials: sum r.z into part_a, max |r| into part_b.
template <bool UPDATE>
__global__ void __launch_bounds__(MDF_THREADS) k_pcg_step(int npad, const int* bperm, const double* binv,
                                                          double* P, const double* Ap, double* Fm, double* R,
                                                          double* Z, const double* px, int npx,
                                                          double* part_a, double* part_b, MdfScalars* Sk) {
  __shared__ double sh[MDF_THREADS / 64];
  __shared__ double rl[MDF_THREADS][3];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = t < npad ? bperm[t] : -1;
  double r[3] = {0.0, 0.0, 0.0};
  if constexpr (UPDATE) {
    if (Sk->done) return;
    const double rz = Sk->rr;
    const double pAp = fold_sum(px, npx, sh);
    const bool ok = pAp > 0.0;
    const double al = ok ? rz / pAp : 0.0;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      Sk->pAp = pAp;  // fields no block of this kernel reads
      Sk->alpha = al;
    }
    if (!ok) return;
    if (k >= 0) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        Fm[3 * (size_t)k + c] += al * P[3 * (size_t)k + c];
        r[c] = R[3 * (size_t)k + c] - al * Ap[3 * (size_t)k + c];
        R[3 * (size_t)k + c] = r[c];
      }
    }
  } else {
    if (k >= 0)
#pragma unroll
      for (int c = 0; c < 3; c++) r[c] = R[3 * (size_t)k + c];
  }
#pragma unroll
  for (int c = 0; c < 3; c++) rl[threadIdx.x][c] = r[c];
  __syncthreads();
  double a = 0.0, b = 0.0;
  if (k >= 0) {
    const int s0 = threadIdx.x & ~(BJ - 1);  // first lane of the block
    // column u of the block's inverse holds row s of it at u * BJ + s
    const double* col = binv + (size_t)(t - (threadIdx.x - s0)) * BJ + (threadIdx.x - s0);
    double z0 = 0.0, z1 = 0.0, z2 = 0.0;
#pragma unroll 8
    for (int u = 0; u < BJ; u++) {
      const double w = col[u * BJ];
      z0 += w * rl[s0 + u][0];
      z1 += w * rl[s0 + u][1];
      z2 += w * rl[s0 + u][2];
    }
    Z[3 * (size_t)k + 0] = z0;
    Z[3 * (size_t)k + 1] = z1;
    Z[3 * (size_t)k + 2] = z2;
    if constexpr (!UPDATE) {
      P[3 * (size_t)k + 0] = z0;
      P[3 * (size_t)k + 1] = z1;
      P[3 * (size_t)k + 2] = z2;
    }
    a = r[0] * z0 + r[1] * z1 + r[2] * z2;
    b = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
  }
  a = block_reduce(a, 0, sh);
  b = block_reduce(b, 1, sh);
  if (threadIdx.x == 0) {
    part_a[blockIdx.x] = a;
    part_b[blockIdx.x] = b;
  }
}

// one block: r.z of the restart into the scalars
__global__ void __launch_bounds__(MDF_THREADS) k_pcg_rz(const double* part_a, int nblk, MdfScalars* S) {
  __shared__ double sh[MDF_THREADS / 64];
  const double rz = fold_sum(part_a, nblk, sh);
  if (threadIdx.x == 0) S->rr = rz;
}

}  // namespace couplb
