// couplb_hip.hip -- device context and C ABI (include/couplb_hip.h).
#include <cstring>
#include <algorithm>
#include <functional>
#include <mutex>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "../../include/couplb_hip.h"
#include "ibm.hpp"
#include "kernels.hpp"
#include "mdf.hpp"
#include "mdf_matrix.hpp"
#include "mlp.hpp"
#include "spread_pairs.hpp"
#include "spread_rows.hpp"

using namespace couplb;

// Device memory of couplb_diagnostics_noalloc: MACRO_PARTS block partials of
// k_macro_parts (9 doubles each) and the six results of k_macro_reduce,
// 18,496 bytes on every GPU the library runs on, held until the process
// ends.  Host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (scratch_on); the name itself is not a
// device address on the host.
static __device__ __align__(256) double scratch[MACRO_PARTS * 9 + 8];
// populations the 16-bit fixed-point storage clamped to its range
__device__ unsigned long long couplb::i16_saturations = 0;
constexpr size_t SCRATCH_OUT = (size_t)MACRO_PARTS * 9;  // offset of the results

namespace {

// Address of `scratch` on each device, looked up once per device with that
// device current (couplb_create).
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
double* scratch_dev[MAX_DEVICES] = {};

// the current device's scratch address, or null if the lookup failed
double* scratch_on(int dev) {
  if (dev < 0 || dev >= MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_dev[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    scratch_dev[dev] = static_cast<double*>(ptr);
  }
  return scratch_dev[dev];
}

// Kernel timing without host synchronisation: each timed region records a
// pair of events into a ring; a pair is read back (hipEventSynchronize) only
// when its ring entry is reused or when the caller asks for the times.
struct TimerRing {
  static constexpr int R = 256;
  static constexpr int W = 5;  // 0 LB kernel, 1 IBM sub-step, 2 spreading, 3 MDF solve,
                               // 4 velocity gather (inside 1 or 3)
  hipEvent_t a[R] = {}, b[R] = {};
  int which[R] = {};
  bool pending[R] = {};
  int head = 0;
  double total[W] = {0, 0, 0, 0, 0}, last[W] = {0, 0, 0, 0, 0};
  long count[W] = {0, 0, 0, 0, 0};
};

struct SlotHost {
  MarkerRec* rec = nullptr;
  int n = 0;
  int epoch = -1;  // marker ordering the records were built from
};

}  // namespace

// chunks of a host-array sub-step that only recomputes forces
constexpr int HOST_CHUNKS_MAX = 8;

struct couplb_ctx {
  couplb_params p{};
  int D = 3, Q = 19;
  Geo geo{};
  Walls walls{};
  size_t N = 0, elem = 8;
  Layout lay{0, 0};   // population layout (kernels.hpp)
  size_t falloc = 0;  // elements per population buffer
  void* f[2] = {nullptr, nullptr};
  int cur = 0;           // buffer holding the state
  int local = 1;         // state is post-stream (create / import)
  int prev_valid = 0;    // a step since the state was set: f[cur ^ 1] holds its input
  double* rho = nullptr; // collision density (interior layout)
  int rho_fresh = 0;     // density refreshed since the last step (P1)
  int has_moving = 0;
  double base = 0.0;
  float scale = 1.0f;
  int cycle = 0;         // state counter for the u stamps
  size_t fused_nodes = 0;  // grids up to this many nodes: k_steps_fused
  unsigned timer_mask = ~0u;  // timed region kinds (bit per kind)
  hipStream_t cstream = nullptr;  // host-array uploads of the IBM sub-steps
  hipEvent_t up_done = nullptr;   // end of the latest upload on cstream
  hipStream_t dstream = nullptr;  // host-array downloads of chunked sub-steps
  hipEvent_t ev_up[HOST_CHUNKS_MAX] = {}, ev_cmp[HOST_CHUNKS_MAX] = {};  // per chunk
  int* iperm = nullptr;           // sorted position of each atom
  int iperm_epoch = -1;           // ordering iperm belongs to
  int hx_n = -1;                  // markers whose host x, m dx, dm hold (-1: none)
  int host_cycle = -1;            // state of the latest host sub-step
  size_t fused_lds = 0;    // its LDS bytes; 0 when steps do not fuse
  hipStream_t stream = nullptr;
  TimerRing tr;
  std::string err;
  // macroscopic staging
  double* mac = nullptr;      // 4*N
  double* partial = nullptr;  // 9*ny*nz
  double* red = nullptr;      // 6 + 3*ny
  // IBM
  double* u4 = nullptr;   // IBM fluid velocity and density (ux, uy, uz, rho) per node
  int* stamp = nullptr;
  double* F4 = nullptr;  // spread force (fx, fy, fz, unused) per node
  unsigned char *flag = nullptr, *cover = nullptr;
  int cap = 0;
  unsigned *keys = nullptr, *skeys = nullptr;
  int *vals = nullptr, *perm = nullptr;
  int* rowstart_cur = nullptr;
  Base* bases = nullptr;
  void* sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  double *dx = nullptr, *dv = nullptr, *dm = nullptr, *df = nullptr;
  int order_n = -1;       // marker count of the cached ordering
  double order_key[5] = {0, 0, 0, 0, -1};  // lo, inv_dx, kernel of the cached ordering
  int pos_epoch = 0;      // incremented whenever the ordering is rebuilt
  int cover_epoch = -1;   // ordering the coverage flags were built from
  double* ufc = nullptr;  // interpolated fluid velocity per sorted marker (3*cap)
  int uf_cycle = -1, uf_epoch = -1;  // state and ordering of ufc
  // fixed-marker spreading (spread_pairs.hpp)
  unsigned *pkey = nullptr, *pkey_s = nullptr;
  int *pval = nullptr, *pval_s = nullptr, *pp = nullptr;
  double* pw = nullptr;
  void* ptmp = nullptr;
  size_t ptmp_bytes = 0;
  long pcap = 0, pne = 0, pcap_runs = 0;
  int pair_epoch = -1;       // ordering the pairs (and the node flags) were built for
  int *pheads = nullptr, *phlen = nullptr;  // node runs of the sorted pairs
  unsigned* phnode = nullptr;
  unsigned short* plen = nullptr;  // per node: pair count of its run (0: none)
  int* pseg = nullptr;             // per 64-node segment of an x-row: its first pair
  int runs_long = 0;               // a run of the map is longer than plen holds
  unsigned char* phflag = nullptr;
  int* pcount = nullptr;           // device count of runs
  int pnh = 0;                     // number of runs
  void* seltmp = nullptr;
  size_t seltmp_bytes = 0;
  double* fsrc = nullptr;          // record slot 0's force per sorted marker (3*cap)
  int prev_step_epoch = -1;  // ordering of the single record slot at the previous LB step
  // moving-marker map (spread_pairs.hpp k_pairs_base): the same pair arrays
  // built from the ordering's base cells, the per-step force of every
  // (marker, stencil slot), and the spreading of the records that left their
  // base cell (F4 where flag2 is set)
  int mpair_epoch = -1;      // ordering the moving-marker map was built for
  double* pcon = nullptr;
  long pcon_cap = 0;
  unsigned char* flag2 = nullptr;
  int f2_tag = 0;                  // tag of the current step's flag2 marks
  unsigned char* tflag = nullptr;  // tiles reached by records that left their cell
  int* tlist = nullptr;            // their list, then its length
  int ntiles = 0;
  int order_age = 0;         // LB intervals since the marker ordering was built
  int reorder_every = 32;    // moving markers keep an ordering at most this many intervals
  // records off the kept ordering's map at the last LB step, reported by the
  // device into pinned memory and read once its event has completed
  int* moved_host = nullptr;
  hipEvent_t moved_ev = nullptr;
  int moved_pending = 0;
  int moved_epoch = -1;            // ordering the pending report belongs to
  long moved_cnt = 0, moved_total = 0;
  int kept_move = 0;         // positions changed on the kept ordering since it was built
  double* vcache = nullptr;  // stencil velocities of moving markers (k_marker VelCache)
  int4* vbase = nullptr;
  int vcache_cap = 0;
  // multi-direct forcing (mdf.hpp)
  double* ut = nullptr;            // operator node values at pair nodes, 3 per node
  double* mdf_v = nullptr;         // target velocity per sorted marker (3*cap)
  double *mdf_F = nullptr, *mdf_R = nullptr, *mdf_Ap = nullptr;  // force, residual, A p
  double* mdf_Z = nullptr;         // preconditioned residual (COUPLB_MDF_PCG)
  double* mdf_part = nullptr;      // block partials
  int mdf_cap = 0;
  MdfScalars* mdf_S = nullptr;
  MdfScalars* mdf_host = nullptr;  // pinned, two slots: chunks of passes alternate
  hipEvent_t mdf_ev[2] = {nullptr, nullptr};  // end of the last two queued chunks
  float *mlp_x = nullptr, *mlp_y = nullptr;  // network inputs / outputs, sorted order
  double* mlp_sig = nullptr;                 // |slip at F = 0| per sorted marker
  // explicit marker-by-marker MDF operator (mdf_matrix.hpp)
  double* mm_gam = nullptr;  // interpolation / spreading weight ratio per sorted marker
  double* mm_B = nullptr;    // b = V - I(u*) per sorted marker
  int mm_epoch = -1;         // pair map the matrix was built for (-1: none)
  int mm_seen = -1;          // pair map of the previous MDF solve
  bool mm_ok = false;        // built (false: over the size limit, node path)
  int mm_n = 0, mm_nnz = 0;
  long long mm_T = 0;        // triples
  unsigned long long* mm_val = nullptr;  // sorted triples: pair of k << 32 | pair of j
  double* mm_q = nullptr;                // per pair: weight / (2 rho) of the current state
  long mm_q_cap = 0;
  int *mm_ent = nullptr, *mm_elen = nullptr, *mm_erow = nullptr, *mm_col = nullptr;
  int* mm_rowptr = nullptr;
  double* mm_aval = nullptr;  // entries for the current density
  // block-Jacobi preconditioner of the explicit operator (mdf_matrix.hpp)
  bool bj_ok = false;         // blocks built for the current matrix
  int bj_npad = 0;            // markers rounded up to whole blocks
  int *bj_perm = nullptr, *bj_pos = nullptr;  // block order <-> sorted marker
  double* bj_inv = nullptr;   // block inverses, BJ x BJ each
  int ugather_cycle = -1; // state counter of the last velocity gather
  std::vector<SlotHost> slots;
  int nslots = 0;
  int f4_partial = 0;  // F4 / flag already hold the spread records of earlier
                       // sub-steps of this interval (slots were full)
  // binned records of the pending slots (spread_rows.hpp)
  int *sb_cnt = nullptr, *sb_start = nullptr;   // bucket counts / starts, nb + 1 each
  int *sb_bk = nullptr, *sb_rank = nullptr;     // bucket and arrival rank per record
  unsigned long long *sb_ent = nullptr, *sb_ent2 = nullptr;  // binned records (spread_rows.hpp entry_of)
  int* sb_large = nullptr;                      // buckets sorted by k_bucket_sort_large, + count
  void* sb_scan_tmp = nullptr;
  size_t sb_scan_bytes = 0;
  long sb_cap = 0;                              // records the per-record arrays hold
  int sb_nb = 0;                                // buckets
  double* staging = nullptr;  // ghost-inclusive FP64 state for import/export
  size_t staging_len = 0;
  double* scratch = nullptr;  // this device's `scratch` (couplb_diagnostics_noalloc)
};

namespace {

int fail(couplb_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                            \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(ctx, COUPLB_EDEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

int timer_harvest(couplb_ctx* c, int e) {
  TimerRing& t = c->tr;
  if (!t.pending[e]) return COUPLB_OK;
  HIPCHK(c, hipEventSynchronize(t.b[e]));
  float ms = 0;
  HIPCHK(c, hipEventElapsedTime(&ms, t.a[e], t.b[e]));
  t.total[t.which[e]] += ms;
  t.count[t.which[e]] += 1;
  t.last[t.which[e]] = ms;
  t.pending[e] = false;
  return COUPLB_OK;
}
int timer_harvest_all(couplb_ctx* c) {
  // oldest first, so `last` ends with the newest region of each kind
  for (int k = 0; k < TimerRing::R; k++) {
    const int rc = timer_harvest(c, (c->tr.head + k) % TimerRing::R);
    if (rc) return rc;
  }
  return COUPLB_OK;
}
// opens a timed region of kind `which`; returns its ring entry in *slot
int timer_begin(couplb_ctx* c, int which, int* slot) {
  if (!((c->timer_mask >> which) & 1u)) {  // kind not timed: no events
    *slot = -1;
    return COUPLB_OK;
  }
  TimerRing& t = c->tr;
  const int e = t.head;
  int rc = timer_harvest(c, e);
  if (rc) return rc;
  t.head = (t.head + 1) % TimerRing::R;
  t.which[e] = which;
  HIPCHK(c, hipEventRecord(t.a[e], c->stream));
  *slot = e;
  return COUPLB_OK;
}
int timer_end(couplb_ctx* c, int slot) {
  if (slot < 0) return COUPLB_OK;
  HIPCHK(c, hipEventRecord(c->tr.b[slot], c->stream));
  c->tr.pending[slot] = true;
  return COUPLB_OK;
}

size_t round_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

int storage_bytes(int s) { return s == COUPLB_FP64 ? 8 : s == COUPLB_FP32 ? 4 : 2; }

// scale of the stored deviation: 1024 for FP16 (lifts small deviations out
// of its subnormal range), for the 16-bit fixed point the largest power of
// two with 32767 * (1 / scale) >= range; 1 otherwise
float storage_scale(const couplb_params* p) {
  if (p->storage == COUPLB_FP16) return 1024.0f;
  if (p->storage != COUPLB_I16) return 1.0f;
  const double range = p->storage_range > 0 ? p->storage_range : 1.0 / 32.0;
  double sc = 1.0;
  while (32767.0 / (2.0 * sc) >= range && sc < 1e30) sc *= 2.0;
  while (32767.0 / sc < range && sc > 1e-30) sc *= 0.5;
  return (float)sc;
}

dim3 row_block(int nx) {
  const int bx = nx > 128 ? 256 : (nx > 64 ? 128 : 64);
  return dim3(bx, 1, 1);
}
dim3 row_grid(const Geo& g, int bx) {
  return dim3((unsigned)((g.nx + bx - 1) / bx), (unsigned)g.ny, (unsigned)(g.D == 3 ? g.nz : 1));
}

PullCtx make_pull(couplb_ctx* c) {
  PullCtx pc;
  pc.src = c->f[c->cur];
  pc.lay = c->lay;
  pc.g = c->geo;
  pc.w = c->walls;
  pc.rho_coll = c->rho;
  pc.rho_floor = std::max(0.1 * c->p.rho0, RHO_MIN);  // couplb_streaming.h:86
  pc.local = c->local;
  pc.base = (c->p.storage == COUPLB_FP64) ? 0.0 : c->p.rho0;
  pc.rho0 = c->p.rho0;
  pc.scale = c->scale;
  pc.inv_scale = 1.0f / c->scale;
  return pc;
}

// Dispatch over lattice x storage.
template <template <int, typename> class F, typename... Args>
void dispatch(couplb_ctx* c, Args&&... args) {
  const int s = c->p.storage;
  if (c->D == 3) {
    if (s == COUPLB_FP64) F<3, double>::run(c, args...);
    else if (s == COUPLB_FP32) F<3, float>::run(c, args...);
    else if (s == COUPLB_FP16) F<3, _Float16>::run(c, args...);
    else if (s == COUPLB_BF16) F<3, bf16_t>::run(c, args...);
    else F<3, i16_t>::run(c, args...);
  } else {
    if (s == COUPLB_FP64) F<2, double>::run(c, args...);
    else if (s == COUPLB_FP32) F<2, float>::run(c, args...);
    else if (s == COUPLB_FP16) F<2, _Float16>::run(c, args...);
    else if (s == COUPLB_BF16) F<2, bf16_t>::run(c, args...);
    else F<2, i16_t>::run(c, args...);
  }
}

template <int D, typename S> struct LaunchStep {
  static void run(couplb_ctx* c, const StepArgs& a) {
    if constexpr (Store<S>::deviation) {
      constexpr int G = LDS_GB / (int)sizeof(S), XC = lds_chunk<S>();
      if (!a.pc.local && c->geo.nx % G == 0 && c->lay.pq % G == 0 && c->lay.pr % G == 0) {
        // reduced formats: 16-byte staged row groups (k_step_lds)
        const dim3 grid((unsigned)((c->geo.nx + XC - 1) / XC), (unsigned)c->geo.ny,
                        (unsigned)(c->D == 3 ? c->geo.nz : 1));
        if (a.pair_mode == 1) hipLaunchKernelGGL((k_step_lds<D, S, 1>), grid, dim3(XC), 0, c->stream, a);
        else if (a.pair_mode == 2) hipLaunchKernelGGL((k_step_lds<D, S, 2>), grid, dim3(XC), 0, c->stream, a);
        else hipLaunchKernelGGL((k_step_lds<D, S>), grid, dim3(XC), 0, c->stream, a);
        return;
      }
    }
    const dim3 b = row_block(c->geo.nx);
    if (a.pair_mode == 1) hipLaunchKernelGGL((k_step<D, S, 1>), row_grid(c->geo, b.x), b, 0, c->stream, a);
    else if (a.pair_mode == 2) hipLaunchKernelGGL((k_step<D, S, 2>), row_grid(c->geo, b.x), b, 0, c->stream, a);
    else hipLaunchKernelGGL((k_step<D, S>), row_grid(c->geo, b.x), b, 0, c->stream, a);
  }
};
template <int D, typename S> struct LaunchMacro {
  static void run(couplb_ctx* c, const MacroArgs& a) {
    const dim3 b = row_block(c->geo.nx);
    hipLaunchKernelGGL((k_macro<D, S>), dim3(1, c->geo.ny, c->D == 3 ? c->geo.nz : 1), b, 0,
                       c->stream, a);
  }
};
template <int D, typename S> struct LaunchFixup {
  static void run(couplb_ctx* c, int n, const double* x, const IbmParams& ip, const PullCtx& pc) {
    const int tb = 256;  // one lane per marker
    hipLaunchKernelGGL((k_moved_fix<D, S>), dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, c->stream, n,
                       c->perm, x, ip, pc, (const Base*)c->bases, c->u4, c->stamp, c->cycle);
  }
};
template <int D, typename S> struct LaunchMacroParts {
  static void run(couplb_ctx* c, const MacroArgs& a) {
    const dim3 b = row_block(c->geo.nx);  // k_macro's block: the same row sums
    hipLaunchKernelGGL((k_macro_parts<D, S>), dim3(MACRO_PARTS), b, 0, c->stream, a);
  }
};
template <int D, typename S> struct LaunchImport {
  static void run(couplb_ctx* c, const double* fin, void* dst) {
    const dim3 b = row_block(c->geo.nx);
    hipLaunchKernelGGL((k_import<D, S>), row_grid(c->geo, b.x), b, 0, c->stream, fin, dst,
                       c->lay, c->geo, c->p.rho0, c->scale);
  }
};
template <int D, typename S> struct LaunchInitEq {
  static void run(couplb_ctx* c, void* dst, double r, const double* u) {
    const dim3 b = row_block(c->geo.nx);
    hipLaunchKernelGGL((k_init_eq<D, S>), row_grid(c->geo, b.x), b, 0, c->stream, dst, c->lay,
                       c->geo, r, u[0], u[1], u[2], c->base, c->scale);
  }
};
template <int D, typename S> struct LaunchExport {
  static void run(couplb_ctx* c, const PullCtx& pc, const void* prev, double* fout) {
    const int gx = c->geo.nx + 2, gy = c->geo.ny + 2, gz = (D == 3) ? c->geo.nz + 2 : 1;
    const dim3 b = row_block(gx);
    const dim3 grid((unsigned)((gx + b.x - 1) / b.x), (unsigned)gy, (unsigned)gz);
    hipLaunchKernelGGL((k_export<D, S>), grid, b, 0, c->stream, pc, prev, fout);
  }
};
template <int D, typename S> struct LaunchUgather {
  static void run(couplb_ctx* c, const PullCtx& pc, const BaseSlot& bs, bool cover) {
    if (cover) {
      const int tb = 256;
      hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks(c->N)), dim3(256), 0, c->stream, c->cover,
                         c->N, (unsigned char)0);
      const long ne = (long)c->order_n * stencil_size<D>();
      if (ne > 0)
        hipLaunchKernelGGL(k_cover_scatter<D>, dim3((unsigned)((ne + tb - 1) / tb)), dim3(tb), 0, c->stream,
                           c->order_n, c->geo, c->walls, bs.rec, c->cover);
    }
    const dim3 b = row_block(c->geo.nx);
    // reduced formats: dense pull when the current marker cells cover more
    // than a quarter of the nodes (the covered-node count of the fixed-marker
    // map, when it belongs to this ordering); measured faster for 16-bit
    // storage and slower for FP64
    const bool dense = sizeof(S) < 8 && (c->pair_epoch == c->pos_epoch || c->mpair_epoch == c->pos_epoch) &&
                       4 * (size_t)c->pnh > c->N;
    constexpr int G = LDS_GB / (int)sizeof(S), XC = lds_chunk<S>();
    const bool staged = !pc.local && c->geo.nx % G == 0 && c->lay.pq % G == 0 && c->lay.pr % G == 0;
    if (dense && staged) {
      if constexpr (Store<S>::deviation) {
        const dim3 grid((unsigned)((c->geo.nx + XC - 1) / XC), (unsigned)c->geo.ny,
                        (unsigned)(D == 3 ? c->geo.nz : 1));
        hipLaunchKernelGGL((k_upull_lds<D, S>), grid, dim3(XC), 0, c->stream, pc, c->cover, c->u4, c->stamp,
                           c->cycle);
      }
    } else if (dense)
      hipLaunchKernelGGL((k_upull<D, S, true>), row_grid(c->geo, b.x), b, 0, c->stream, pc, c->cover, c->u4,
                         c->stamp, c->cycle, 1);
    else
      hipLaunchKernelGGL((k_upull<D, S, false>), row_grid(c->geo, b.x), b, 0, c->stream, pc, c->cover,
                         c->u4, c->stamp, c->cycle, c->ugather_cycle != c->cycle ? 1 : 0);
  }
};

int ensure_macro_buffers(couplb_ctx* c) {
  if (!c->mac) HIPCHK(c, hipMalloc(&c->mac, sizeof(double) * 4 * c->N));
  if (!c->partial)
    HIPCHK(c, hipMalloc(&c->partial, sizeof(double) * 9 * (size_t)c->geo.ny * std::max(1, c->geo.nz)));
  if (!c->red) HIPCHK(c, hipMalloc(&c->red, sizeof(double) * (6 + 3 * (size_t)c->geo.ny)));
  return COUPLB_OK;
}

int run_macro(couplb_ctx* c, bool want_fields) {
  int rc = ensure_macro_buffers(c);
  if (rc) return rc;
  MacroArgs a;
  a.pc = make_pull(c);
  a.rho = want_fields ? c->mac : nullptr;
  a.ux = want_fields ? c->mac + c->N : nullptr;
  a.uy = want_fields ? c->mac + 2 * c->N : nullptr;
  a.uz = want_fields ? c->mac + 3 * c->N : nullptr;
  a.partial = c->partial;
  dispatch<LaunchMacro>(c, a);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_macro_reduce, dim3(1), dim3(256), 0, c->stream, c->partial, c->geo.ny,
                     c->D == 3 ? c->geo.nz : 1, c->red, c->red + 6);
  HIPCHK(c, hipGetLastError());
  c->rho_fresh = 1;
  return COUPLB_OK;
}

unsigned key_bits(size_t nrows) {
  unsigned b = 1;
  while (b < 32 && ((size_t)1 << b) <= nrows) b++;
  return b;
}

int ensure_ibm(couplb_ctx* c, int n) {
  if (!c->u4) {
    HIPCHK(c, hipMalloc(&c->u4, sizeof(double) * 4 * c->N));
    HIPCHK(c, hipMalloc(&c->stamp, sizeof(int) * c->N));
    HIPCHK(c, hipMemsetAsync(c->stamp, 0xFF, sizeof(int) * c->N, c->stream));
    HIPCHK(c, hipMalloc(&c->rowstart_cur, sizeof(int) * (c->geo.nrows + 1)));
    HIPCHK(c, hipMalloc(&c->F4, sizeof(double) * 4 * c->N));
    HIPCHK(c, hipMalloc(&c->flag, c->N));
    HIPCHK(c, hipMalloc(&c->cover, c->N));
  }
  if (n <= c->cap) return COUPLB_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int cap = std::max(n, 2 * c->cap);
  (void)hipFree(c->keys); (void)hipFree(c->skeys); (void)hipFree(c->vals); (void)hipFree(c->perm); (void)hipFree(c->bases);
  (void)hipFree(c->sort_tmp); (void)hipFree(c->dx); (void)hipFree(c->dv); (void)hipFree(c->dm); (void)hipFree(c->df);
  (void)hipFree(c->ufc); (void)hipFree(c->iperm);
  c->iperm = nullptr;
  c->iperm_epoch = -1;
  HIPCHK(c, hipMalloc(&c->keys, sizeof(unsigned) * cap));
  HIPCHK(c, hipMalloc(&c->skeys, sizeof(unsigned) * cap));
  HIPCHK(c, hipMalloc(&c->vals, sizeof(int) * cap));
  HIPCHK(c, hipMalloc(&c->perm, sizeof(int) * cap));
  HIPCHK(c, hipMalloc(&c->bases, sizeof(Base) * cap));
  HIPCHK(c, hipMalloc(&c->dx, sizeof(double) * 3 * cap));
  HIPCHK(c, hipMalloc(&c->dv, sizeof(double) * 3 * cap));
  HIPCHK(c, hipMalloc(&c->dm, sizeof(double) * cap));
  HIPCHK(c, hipMalloc(&c->df, sizeof(double) * 3 * cap));
  HIPCHK(c, hipMalloc(&c->ufc, sizeof(double) * 3 * cap));
  HIPCHK(c, hipMalloc(&c->iperm, sizeof(int) * cap));
  (void)hipFree(c->fsrc);
  c->fsrc = nullptr;
  HIPCHK(c, hipMalloc(&c->fsrc, sizeof(double) * 3 * cap));
  (void)hipFree(c->mdf_v);
  c->mdf_v = nullptr;
  HIPCHK(c, hipMalloc(&c->mdf_v, sizeof(double) * 3 * cap));
  size_t bytes = 0;
  HIPCHK(c, rocprim::radix_sort_pairs(nullptr, bytes, c->keys, c->skeys, c->vals, c->perm,
                                      (size_t)cap, 0u, key_bits(c->geo.nrows), c->stream));
  c->sort_tmp_bytes = bytes;
  HIPCHK(c, hipMalloc(&c->sort_tmp, bytes));
  // record slots: the pending ones (sub-steps of this LB interval not yet
  // spread) keep their records, so a group that grows within an interval
  // still spreads what the earlier sub-steps recorded
  for (int s = 0; s < (int)c->slots.size(); s++) {
    SlotHost& sl = c->slots[s];
    MarkerRec* nr = nullptr;
    HIPCHK(c, hipMalloc(&nr, sizeof(MarkerRec) * cap));
    if (s < c->nslots && sl.n > 0)
      HIPCHK(c, hipMemcpyAsync(nr, sl.rec, sizeof(MarkerRec) * (size_t)sl.n, hipMemcpyDeviceToDevice,
                               c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(sl.rec);
    sl.rec = nr;
  }
  c->cap = cap;
  c->hx_n = -1;
  c->order_n = -1;
  c->uf_epoch = -1;
  return COUPLB_OK;
}

int ensure_slot(couplb_ctx* c, int s) {
  while ((int)c->slots.size() <= s) {
    if ((int)c->slots.size() >= MAX_SLOTS)
      return fail(c, COUPLB_ESTATE, "more than 16 IBM sub-steps between two LB steps");
    SlotHost h;
    HIPCHK(c, hipMalloc(&h.rec, sizeof(MarkerRec) * std::max(1, c->cap)));
    c->slots.push_back(h);
  }
  return COUPLB_OK;
}

IbmParams make_ibm_params(const couplb_ctx* c, const couplb_ibm_args* a) {
  IbmParams ip;
  for (int d = 0; d < 3; d++) ip.lo[d] = a->lo[d];
  ip.inv_dx = a->inv_dx;
  ip.xi_over_dtLB = a->xi_over_dtLB;
  ip.vel_scale = a->vel_scale;
  ip.inv_N = a->inv_N;
  ip.force_scale = a->force_scale;
  ip.kern = c->p.kernel;
  return ip;
}

// Marker ordering by base row (kept while the caller declares the positions
// unchanged, and over the later sub-steps of an LB interval: keep_order)
// and the fluid velocity at covered nodes of the current state (once per
// state).  Sets *reorder when the ordering was rebuilt.
int prepare_markers(couplb_ctx* c, int n, const double* x, const IbmParams& ip,
                    int positions_unchanged, bool keep_order, bool* reorder_out) {
  const int tb = 256, nb = (n + tb - 1) / tb;
  const double key[5] = {ip.lo[0], ip.lo[1], ip.lo[2], ip.inv_dx, (double)c->p.kernel};
  const bool reorder = !((positions_unchanged || keep_order) && c->order_n == n &&
                         std::memcmp(key, c->order_key, sizeof(key)) == 0);
  if (reorder) {
    hipLaunchKernelGGL(k_keys, dim3(nb), dim3(tb), 0, c->stream, n, x, ip, c->geo, c->keys, c->vals);
    HIPCHK(c, hipGetLastError());
    size_t bytes = c->sort_tmp_bytes;
    HIPCHK(c, rocprim::radix_sort_pairs(c->sort_tmp, bytes, c->keys, c->skeys, c->vals, c->perm,
                                        (size_t)n, 0u, key_bits(c->geo.nrows), c->stream));
    hipLaunchKernelGGL(k_rowstart, dim3((c->geo.nrows + 1 + tb - 1) / tb), dim3(tb), 0, c->stream,
                       c->skeys, n, c->geo.nrows, c->rowstart_cur);
    hipLaunchKernelGGL(k_bases, dim3(nb), dim3(tb), 0, c->stream, n, c->perm, x, ip, c->geo, c->bases);
    HIPCHK(c, hipGetLastError());
    c->order_n = n;
    std::memcpy(c->order_key, key, sizeof(key));
    c->pos_epoch++;
    c->order_age = 0;
    c->kept_move = 0;
    c->moved_cnt = 0;    // counts of the previous ordering no longer apply
    c->moved_total = 0;
  } else if (!positions_unchanged) {
    c->kept_move = 1;  // positions changed on the kept ordering
  }
  if (reorder || c->ugather_cycle != c->cycle) {
    // coverage depends on the marker cells only, the velocities on the state
    BaseSlot bs{c->rowstart_cur, c->bases};
    PullCtx pc = make_pull(c);
    int tg = -1;
    int rc = timer_begin(c, 4, &tg);
    if (rc) return rc;
    dispatch<LaunchUgather>(c, pc, bs, c->cover_epoch != c->pos_epoch);
    HIPCHK(c, hipGetLastError());
    rc = timer_end(c, tg);
    if (rc) return rc;
    c->cover_epoch = c->pos_epoch;
    c->ugather_cycle = c->cycle;
  }
  if (!reorder && !positions_unchanged) {
    // moved markers on the kept ordering: velocities of nodes they newly reach
    PullCtx pc = make_pull(c);
    dispatch<LaunchFixup>(c, n, x, ip, pc);
    HIPCHK(c, hipGetLastError());
  }
  *reorder_out = reorder;
  return COUPLB_OK;
}

int fold_slots(couplb_ctx* c);

// Same positions and ordering, and the same LB state as the previous
// sub-step of this interval: u_f is known and only the force is new
// (k_marker_force); prepare_markers then has nothing to do.
bool force_only(const couplb_ctx* c, int n, int positions_unchanged, const IbmParams& ip) {
  const double key[5] = {ip.lo[0], ip.lo[1], ip.lo[2], ip.inv_dx, (double)c->p.kernel};
  return positions_unchanged && c->nslots > 0 && c->order_n == n &&
         std::memcmp(key, c->order_key, sizeof(key)) == 0 &&
         c->slots[c->nslots - 1].epoch == c->pos_epoch && c->slots[c->nslots - 1].n == n &&
         c->uf_cycle == c->cycle && c->uf_epoch == c->pos_epoch;
}

int ibm_impl(couplb_ctx* c, int n, const double* x, const double* v, const double* m,
             double* f_dev, int fmode, const couplb_ibm_args* a) {
  if (n < 0 || !a) return fail(c, COUPLB_EINVAL, "bad IBM arguments");
  // binned records carry the marker index in 24 bits (spread_rows.hpp)
  if (n >= (1 << REC_SLOT_SHIFT)) return fail(c, COUPLB_EINVAL, "an IBM group holds at most 16777215 atoms");
  if (n == 0) return COUPLB_OK;
  int rc = ensure_ibm(c, n);
  if (rc) return rc;
  const IbmParams ip = make_ibm_params(c, a);
  const int tb = 256, nb = (n + tb - 1) / tb;
  int tslot = 0;
  rc = timer_begin(c, 1, &tslot);
  if (rc) return rc;
  bool reorder = false;
  // a later sub-step of the interval keeps the interval's marker ordering;
  // moving markers keep it over reorder_every intervals once the LB step has
  // built its moving-marker map (markers that leave their base cell are
  // spread apart until the next ordering)
  const bool later = c->nslots > 0 || c->f4_partial;
  if (c->moved_pending) {
    const hipError_t q = hipEventQuery(c->moved_ev);
    if (q == hipSuccess) {
      if (c->moved_epoch == c->pos_epoch) c->moved_cnt = *c->moved_host;
      c->moved_pending = 0;
    } else if (q == hipErrorNotReady) {
      (void)hipGetLastError();  // not an error: the report is still on its way
    } else {
      return fail(c, COUPLB_EDEVICE, hipGetErrorString(q));
    }
  }
  // At the start of an interval a kept ordering that markers moved on is
  // renewed when it is old, when more than a quarter of the last LB step's
  // records had left its base cells, or when the positions are declared
  // unchanged (the markers stopped: the fixed-marker path applies again).
  const bool drifted = c->kept_move && 4 * c->moved_cnt > c->moved_total;
  const bool renew = !later && c->kept_move &&
                     (c->order_age >= c->reorder_every || drifted || a->positions_unchanged);
  const bool keep = later || (c->mpair_epoch == c->pos_epoch && c->order_age < c->reorder_every && !renew);
  const int unchanged = a->positions_unchanged && !renew;
  rc = prepare_markers(c, n, x, ip, unchanged, keep, &reorder);
  if (rc) return rc;
  // Same positions and same LB state as the previous sub-step of this
  // interval: u_f is known and only the force is new (k_marker_force).
  const bool merge = !reorder && a->positions_unchanged && c->nslots > 0 &&
                     c->slots[c->nslots - 1].epoch == c->pos_epoch &&
                     c->slots[c->nslots - 1].n == n && c->uf_cycle == c->cycle &&
                     c->uf_epoch == c->pos_epoch;
  if (merge) {
    SlotHost& sl = c->slots[c->nslots - 1];
    hipLaunchKernelGGL(k_marker_force, dim3(nb), dim3(tb), 0, c->stream, n, c->perm, v, m, f_dev, fmode,
                       ip, (const double*)c->ufc, sl.rec, c->nslots == 1 ? c->fsrc : nullptr);
    HIPCHK(c, hipGetLastError());
    c->rho_fresh = 1;  // compute_macroscopic(false) at fix_couplb.cpp:758
    return timer_end(c, tslot);
  }
  if (c->nslots == MAX_SLOTS) {  // every slot pending: spread them into F4 first
    rc = fold_slots(c);
    if (rc) return rc;
  }
  const int s = c->nslots;
  rc = ensure_slot(c, s);
  if (rc) return rc;
  if (c->slots[s].rec == nullptr || c->cap < n) return fail(c, COUPLB_ENOMEM, "slot capacity");
  SlotHost& sl = c->slots[s];
  // markers that move on a kept ordering: stencil velocities cached for the
  // later sub-steps of the state (k_marker VelCache)
  VelCache vc{nullptr, nullptr, 0, c->cycle};
  if (c->kept_move) {
    const int lpm = (c->D == 3) ? 4 : 1;
    if (c->vcache_cap < c->cap) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      (void)hipFree(c->vcache); (void)hipFree(c->vbase);
      c->vcache = nullptr; c->vbase = nullptr; c->vcache_cap = 0;
      HIPCHK(c, hipMalloc(&c->vcache, sizeof(double) * 48 * lpm * (size_t)c->cap));
      HIPCHK(c, hipMalloc(&c->vbase, sizeof(int4) * (size_t)c->cap));
      // no base cell matches state -1
      HIPCHK(c, hipMemsetAsync(c->vbase, 0xFF, sizeof(int4) * (size_t)c->cap, c->stream));
      c->vcache_cap = c->cap;
    }
    vc.u = c->vcache;
    vc.base = c->vbase;
    vc.L = (long)lpm * c->vcache_cap;
  }
  if (c->D == 3) {
    const int nb4 = (4 * n + tb - 1) / tb;  // four lanes per marker
    hipLaunchKernelGGL((k_marker<3>), dim3(nb4), dim3(tb), 0, c->stream, n, c->perm, x, v, m, f_dev,
                       fmode, ip, c->geo, c->walls, c->u4, sl.rec, c->ufc, s == 0 ? c->fsrc : nullptr, vc);
  } else {
    hipLaunchKernelGGL((k_marker<2>), dim3(nb), dim3(tb), 0, c->stream, n, c->perm, x, v, m, f_dev,
                       fmode, ip, c->geo, c->walls, c->u4, sl.rec, c->ufc, s == 0 ? c->fsrc : nullptr, vc);
  }
  HIPCHK(c, hipGetLastError());
  rc = timer_end(c, tslot);
  if (rc) return rc;
  sl.n = n;
  sl.epoch = c->pos_epoch;
  c->uf_cycle = c->cycle;
  c->uf_epoch = c->pos_epoch;
  c->nslots = s + 1;
  c->rho_fresh = 1;  // compute_macroscopic(false) at fix_couplb.cpp:758
  return COUPLB_OK;
}

// Pair arrays of ne (marker, stencil slot) pairs: keys, values and the
// sort's scratch.
int ensure_pairs(couplb_ctx* c, long ne) {
  if (ne <= c->pcap) return COUPLB_OK;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(c->pkey); (void)hipFree(c->pkey_s); (void)hipFree(c->pval); (void)hipFree(c->pval_s);
  (void)hipFree(c->pp); (void)hipFree(c->pw); (void)hipFree(c->ptmp);
  c->pkey = c->pkey_s = nullptr; c->pval = c->pval_s = c->pp = nullptr; c->pw = nullptr; c->ptmp = nullptr;
  c->pcap = 0;
  HIPCHK(c, hipMalloc(&c->pkey, sizeof(unsigned) * ne));
  HIPCHK(c, hipMalloc(&c->pkey_s, sizeof(unsigned) * ne));
  HIPCHK(c, hipMalloc(&c->pval, sizeof(int) * ne));
  HIPCHK(c, hipMalloc(&c->pval_s, sizeof(int) * ne));
  HIPCHK(c, hipMalloc(&c->pp, sizeof(int) * ne));
  HIPCHK(c, hipMalloc(&c->pw, sizeof(double) * ne));
  size_t bytes = 0;
  HIPCHK(c, rocprim::radix_sort_pairs(nullptr, bytes, c->pkey, c->pkey_s, c->pval, c->pval_s, (size_t)ne, 0u,
                                      key_bits(c->N), c->stream));
  HIPCHK(c, hipMalloc(&c->ptmp, bytes));
  c->ptmp_bytes = bytes;
  c->pcap = ne;
  return COUPLB_OK;
}

// Node runs of the sorted pairs: first pair of each run, its length and
// node; for the LB step the run length of every node and the first pair of
// every 64-node segment of an x-row (kernels.hpp pair_run).  Sets runs_long
// where a node has more pairs than plen holds (then the map is not used).
int build_runs(couplb_ctx* c, long ne) {
  const int tb = 256;
  const unsigned nb = (unsigned)((ne + tb - 1) / tb);
  const unsigned sentinel = (unsigned)c->N;
  if (ne > c->pcap_runs) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->pheads); (void)hipFree(c->phlen); (void)hipFree(c->phnode); (void)hipFree(c->phflag);
    (void)hipFree(c->seltmp);
    c->pheads = c->phlen = nullptr; c->phnode = nullptr; c->phflag = nullptr; c->seltmp = nullptr;
    c->pcap_runs = 0;
    HIPCHK(c, hipMalloc(&c->pheads, sizeof(int) * ne));
    HIPCHK(c, hipMalloc(&c->phlen, sizeof(int) * ne));
    HIPCHK(c, hipMalloc(&c->phnode, sizeof(unsigned) * ne));
    HIPCHK(c, hipMalloc(&c->phflag, ne));
    if (!c->pcount) HIPCHK(c, hipMalloc(&c->pcount, 2 * sizeof(int)));  // run count, long-run flag
    size_t bytes = 0;
    HIPCHK(c, rocprim::select(nullptr, bytes, rocprim::counting_iterator<int>(0), c->phflag, c->pheads,
                              c->pcount, (size_t)ne, c->stream));
    HIPCHK(c, hipMalloc(&c->seltmp, bytes));
    c->seltmp_bytes = bytes;
    c->pcap_runs = ne;
  }
  const int nsx = (c->geo.nx + 63) / 64;
  const long nseg = (long)nsx * (c->N / c->geo.nx);
  if (!c->plen) {
    HIPCHK(c, hipMalloc(&c->plen, sizeof(unsigned short) * c->N));
    HIPCHK(c, hipMalloc(&c->pseg, sizeof(int) * nseg));
  }
  hipLaunchKernelGGL(k_pair_heads, dim3(nb), dim3(tb), 0, c->stream, ne, (const unsigned*)c->pkey_s, sentinel,
                     c->phflag);
  size_t sbytes = c->seltmp_bytes;
  HIPCHK(c, rocprim::select(c->seltmp, sbytes, rocprim::counting_iterator<int>(0), c->phflag, c->pheads,
                            c->pcount, (size_t)ne, c->stream));
  hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks(sizeof(unsigned short) * c->N)), dim3(256), 0, c->stream,
                     reinterpret_cast<unsigned char*>(c->plen), sizeof(unsigned short) * c->N, (unsigned char)0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->pcount + 1, (size_t)1, 0);
  // the runs number at most ne: the kernel reads their count on the device
  hipLaunchKernelGGL(k_head_info, dim3(nb), dim3(tb), 0, c->stream, ne, (const int*)c->pcount,
                     (const int*)c->pheads, (const unsigned*)c->pkey_s, ne, c->phlen, c->phnode, c->plen,
                     c->pcount + 1);
  hipLaunchKernelGGL(k_seg_start, dim3((unsigned)((nseg + tb - 1) / tb)), dim3(tb), 0, c->stream, nseg, nsx,
                     c->geo.nx, (const unsigned*)c->pkey_s, ne, c->pseg);
  HIPCHK(c, hipGetLastError());
  int info[2] = {0, 0};
  HIPCHK(c, hipMemcpyAsync(info, c->pcount, sizeof(info), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->pnh = info[0];
  c->runs_long = info[1];
  c->pne = ne;
  return COUPLB_OK;
}

// Sparse spreading map of the single record slot (spread_pairs.hpp); also
// rewrites every node flag.
int build_pairs(couplb_ctx* c, const SlotHost& sl) {
  const int S = (c->D == 3) ? stencil_size<3>() : stencil_size<2>();
  const long ne = (long)sl.n * S;
  const unsigned sentinel = (unsigned)c->N;
  const unsigned bits = key_bits(c->N);
  int rc = ensure_pairs(c, ne);
  if (rc) return rc;
  const int tb = 256;
  const unsigned nb = (unsigned)((ne + tb - 1) / tb);
  if (c->D == 3)
    hipLaunchKernelGGL(k_pairs<3>, dim3(nb), dim3(tb), 0, c->stream, sl.n, c->geo, c->walls, sl.rec,
                       sentinel, c->pkey, c->pval);
  else
    hipLaunchKernelGGL(k_pairs<2>, dim3(nb), dim3(tb), 0, c->stream, sl.n, c->geo, c->walls, sl.rec,
                       sentinel, c->pkey, c->pval);
  HIPCHK(c, hipGetLastError());
  size_t bytes = c->ptmp_bytes;
  HIPCHK(c, rocprim::radix_sort_pairs(c->ptmp, bytes, c->pkey, c->pkey_s, c->pval, c->pval_s,
                                      (size_t)ne, 0u, bits, c->stream));
  hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks(c->N)), dim3(256), 0, c->stream, c->flag,
                     c->N, (unsigned char)0);
  if (c->D == 3)
    hipLaunchKernelGGL(k_pair_weights<3>, dim3(nb), dim3(tb), 0, c->stream, ne, c->pkey_s, c->pval_s,
                       sentinel, sl.rec, c->pp, c->pw, c->flag);
  else
    hipLaunchKernelGGL(k_pair_weights<2>, dim3(nb), dim3(tb), 0, c->stream, ne, c->pkey_s, c->pval_s,
                       sentinel, sl.rec, c->pp, c->pw, c->flag);
  HIPCHK(c, hipGetLastError());
  rc = build_runs(c, ne);
  if (rc) return rc;
  c->pair_epoch = sl.epoch;
  c->mpair_epoch = -1;  // the pair arrays and node flags now hold this map
  return COUPLB_OK;
}

// The moving-marker map of the current ordering (spread_pairs.hpp
// k_pairs_base): pairs from the ordering's base cells and their runs;
// rewrites every node flag.
int build_moving_map(couplb_ctx* c, int n) {
  const int S = (c->D == 3) ? stencil_size<3>() : stencil_size<2>();
  const long ne = (long)n * S;
  const unsigned sentinel = (unsigned)c->N;
  int rc = ensure_pairs(c, ne);
  if (rc) return rc;
  if (ne > c->pcon_cap) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->pcon);
    c->pcon = nullptr;
    c->pcon_cap = 0;
    HIPCHK(c, hipMalloc(&c->pcon, sizeof(double) * 4 * ne));
    c->pcon_cap = ne;
  }
  const int tb = 256;
  const unsigned nb = (unsigned)((ne + tb - 1) / tb);
  if (c->D == 3)
    hipLaunchKernelGGL(k_pairs_base<3>, dim3(nb), dim3(tb), 0, c->stream, n, c->geo, c->walls,
                       (const Base*)c->bases, sentinel, c->pkey, c->pval);
  else
    hipLaunchKernelGGL(k_pairs_base<2>, dim3(nb), dim3(tb), 0, c->stream, n, c->geo, c->walls,
                       (const Base*)c->bases, sentinel, c->pkey, c->pval);
  HIPCHK(c, hipGetLastError());
  size_t bytes = c->ptmp_bytes;
  HIPCHK(c, rocprim::radix_sort_pairs(c->ptmp, bytes, c->pkey, c->pkey_s, c->pval, c->pval_s, (size_t)ne, 0u,
                                      key_bits(c->N), c->stream));
  hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks(c->N)), dim3(256), 0, c->stream, c->flag, c->N,
                     (unsigned char)0);
  hipLaunchKernelGGL(k_map_flags, dim3(nb), dim3(tb), 0, c->stream, ne, (const unsigned*)c->pkey_s, sentinel,
                     c->flag);
  HIPCHK(c, hipGetLastError());
  rc = build_runs(c, ne);
  if (rc) return rc;
  c->mpair_epoch = c->pos_epoch;
  c->pair_epoch = -1;  // the pair arrays and node flags now hold this map
  return COUPLB_OK;
}

void mm_free(couplb_ctx* c) {
  (void)hipFree(c->mm_val); (void)hipFree(c->mm_ent); (void)hipFree(c->mm_elen);
  (void)hipFree(c->mm_erow); (void)hipFree(c->mm_col); (void)hipFree(c->mm_rowptr);
  (void)hipFree(c->mm_aval);
  c->mm_val = nullptr;
  c->mm_ent = c->mm_elen = c->mm_erow = c->mm_col = c->mm_rowptr = nullptr;
  c->mm_aval = nullptr;
  c->mm_ok = false;
  (void)hipFree(c->bj_perm); (void)hipFree(c->bj_pos); (void)hipFree(c->bj_inv);
  c->bj_perm = c->bj_pos = nullptr;
  c->bj_inv = nullptr;
  c->bj_ok = false;
  c->bj_npad = 0;
  c->mm_nnz = 0;
  c->mm_T = 0;
}

// Largest triple count the explicit operator is built for (32 bytes a triple
// while building, 8 kept); beyond it the solve stays on the node path.
constexpr long long MM_MAX_TRIPLES = 1ll << 27;

// Structure of the explicit operator for the current pair map of n sorted
// markers: triples (k, j, pair of k, pair of j) of every node run, sorted by
// (k, j); one matrix entry per distinct (k, j).
int mm_build(couplb_ctx* c, int n) {
  HIPCHK(c, hipStreamSynchronize(c->stream));
  mm_free(c);
  c->mm_epoch = c->pair_epoch;
  c->mm_n = n;
  const int nh = c->pnh;
  if (nh <= 0) return COUPLB_OK;
  struct Tmp {  // build temporaries, freed on every exit
    hipStream_t s;
    std::vector<void*> p;
    ~Tmp() {
      (void)hipStreamSynchronize(s);
      for (void* q : p) (void)hipFree(q);
    }
  } tmp{c->stream, {}};
  auto alloc = [&](void** q, size_t bytes) -> int {
    HIPCHK(c, hipMalloc(q, bytes));
    tmp.p.push_back(*q);
    return COUPLB_OK;
  };
  const int tb = 256;
  unsigned long long *cnt = nullptr, *off = nullptr;
  int rc = alloc((void**)&cnt, 8 * ((size_t)nh + 1));
  if (!rc) rc = alloc((void**)&off, 8 * ((size_t)nh + 1));
  if (rc) return rc;
  hipLaunchKernelGGL(k_mm_count, dim3((unsigned)((nh + tb) / tb)), dim3(tb), 0, c->stream, nh,
                     (const int*)c->phlen, cnt);
  size_t bytes = 0;
  HIPCHK(c, rocprim::exclusive_scan(nullptr, bytes, cnt, off, 0ull, (size_t)nh + 1,
                                    rocprim::plus<unsigned long long>(), c->stream));
  void* scan_tmp = nullptr;
  if ((rc = alloc(&scan_tmp, bytes))) return rc;
  HIPCHK(c, rocprim::exclusive_scan(scan_tmp, bytes, cnt, off, 0ull, (size_t)nh + 1,
                                    rocprim::plus<unsigned long long>(), c->stream));
  unsigned long long T = 0;
  HIPCHK(c, hipMemcpyAsync(&T, off + nh, sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (T == 0 || T > (unsigned long long)MM_MAX_TRIPLES) return COUPLB_OK;
  // the build holds about 40 bytes per triple (keys, values, sorted copies,
  // flags, entry starts, sort scratch); leave half of the free memory alone
  size_t mem_free = 0, mem_total = 0;
  HIPCHK(c, hipMemGetInfo(&mem_free, &mem_total));
  if (40ull * T > mem_free / 2) return COUPLB_OK;
  unsigned long long *key = nullptr, *key_s = nullptr, *val = nullptr;
  unsigned char* flag = nullptr;
  int *ent = nullptr, *cnt_d = nullptr;
  if ((rc = alloc((void**)&key, 8 * T)) || (rc = alloc((void**)&key_s, 8 * T)) ||
      (rc = alloc((void**)&val, 8 * T)) || (rc = alloc((void**)&flag, T)) ||
      (rc = alloc((void**)&ent, 4 * T)) || (rc = alloc((void**)&cnt_d, sizeof(int))))
    return rc;
  HIPCHK(c, hipMalloc(&c->mm_val, 8 * T));
  const unsigned nbt = (unsigned)((T + tb - 1) / tb);
  const unsigned long long nn = (unsigned long long)n;
  hipLaunchKernelGGL(k_mm_emit, dim3((unsigned)((nh + tb - 1) / tb)), dim3(tb), 0, c->stream, nh,
                     (const int*)c->pheads, (const int*)c->phlen, (const int*)c->pp, nn,
                     (const unsigned long long*)off, key, val);
  unsigned bits = 1;
  while (bits < 64 && (1ull << bits) < nn * nn) bits++;
  bytes = 0;
  HIPCHK(c, rocprim::radix_sort_pairs(nullptr, bytes, key, key_s, val, c->mm_val, (size_t)T, 0u, bits,
                                      c->stream));
  void* sort_tmp = nullptr;
  if ((rc = alloc(&sort_tmp, bytes))) return rc;
  HIPCHK(c, rocprim::radix_sort_pairs(sort_tmp, bytes, key, key_s, val, c->mm_val, (size_t)T, 0u, bits,
                                      c->stream));
  hipLaunchKernelGGL(k_mm_heads, dim3(nbt), dim3(tb), 0, c->stream, (long long)T,
                     (const unsigned long long*)key_s, flag);
  bytes = 0;
  HIPCHK(c, rocprim::select(nullptr, bytes, rocprim::counting_iterator<int>(0), flag, ent, cnt_d,
                            (size_t)T, c->stream));
  void* sel_tmp = nullptr;
  if ((rc = alloc(&sel_tmp, bytes))) return rc;
  HIPCHK(c, rocprim::select(sel_tmp, bytes, rocprim::counting_iterator<int>(0), flag, ent, cnt_d,
                            (size_t)T, c->stream));
  int nnz = 0;
  HIPCHK(c, hipMemcpyAsync(&nnz, cnt_d, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMalloc(&c->mm_ent, sizeof(int) * (size_t)nnz));
  HIPCHK(c, hipMalloc(&c->mm_elen, sizeof(int) * (size_t)nnz));
  HIPCHK(c, hipMalloc(&c->mm_erow, sizeof(int) * (size_t)nnz));
  HIPCHK(c, hipMalloc(&c->mm_col, sizeof(int) * (size_t)nnz));
  HIPCHK(c, hipMalloc(&c->mm_aval, sizeof(double) * (size_t)nnz));
  HIPCHK(c, hipMalloc(&c->mm_rowptr, sizeof(int) * ((size_t)n + 1)));
  HIPCHK(c, hipMemcpyAsync(c->mm_ent, ent, sizeof(int) * (size_t)nnz, hipMemcpyDeviceToDevice, c->stream));
  hipLaunchKernelGGL(k_mm_entries, dim3((unsigned)((nnz + tb - 1) / tb)), dim3(tb), 0, c->stream, nnz,
                     (const int*)c->mm_ent, (const unsigned long long*)key_s, (long long)T, nn, c->mm_col,
                     c->mm_erow, c->mm_elen);
  hipLaunchKernelGGL(k_mm_rowptr, dim3((unsigned)((n + tb) / tb)), dim3(tb), 0, c->stream, n, nnz,
                     (const int*)c->mm_erow, c->mm_rowptr);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mm_ok = true;
  c->mm_T = (long long)T;
  c->mm_nnz = nnz;
  return COUPLB_OK;
}

// Blocks of the preconditioner for the explicit operator of n sorted markers
// with records rec: Morton order of their base cells, BJ markers a block.
int bj_build(couplb_ctx* c, int n, const MarkerRec* rec) {
  const int tb = 256;
  const int npad = (n + BJ - 1) / BJ * BJ;
  if (c->bj_perm || c->bj_pos || c->bj_inv) {  // left by a build that failed part-way
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->bj_perm); (void)hipFree(c->bj_pos); (void)hipFree(c->bj_inv);
    c->bj_perm = c->bj_pos = nullptr;
    c->bj_inv = nullptr;
  }
  unsigned long long *key = nullptr, *key_s = nullptr;
  int *val = nullptr, *val_s = nullptr;
  void* tmp = nullptr;
  auto release = [&]() {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(key); (void)hipFree(key_s); (void)hipFree(val); (void)hipFree(val_s); (void)hipFree(tmp);
  };
  if (hipMalloc(&key, 8 * (size_t)n) != hipSuccess || hipMalloc(&key_s, 8 * (size_t)n) != hipSuccess ||
      hipMalloc(&val, 4 * (size_t)n) != hipSuccess || hipMalloc(&val_s, 4 * (size_t)n) != hipSuccess) {
    release();
    return fail(c, COUPLB_ENOMEM, "out of device memory for the MDF preconditioner");
  }
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, key, key_s, val, val_s, (size_t)n, 0u, 63u,
                                           c->stream);
  if (e == hipSuccess) e = hipMalloc(&tmp, bytes);
  if (e == hipSuccess) e = hipMalloc(&c->bj_perm, 4 * (size_t)npad);
  if (e == hipSuccess) e = hipMalloc(&c->bj_pos, 4 * (size_t)n);
  if (e == hipSuccess) e = hipMalloc(&c->bj_inv, sizeof(double) * BJ * (size_t)npad);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bj_keys, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, c->stream, n, rec, key,
                       val);
    e = rocprim::radix_sort_pairs(tmp, bytes, key, key_s, val, val_s, (size_t)n, 0u, 63u, c->stream);
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_bj_perm, dim3((unsigned)((npad + tb - 1) / tb)), dim3(tb), 0, c->stream, n, npad,
                       (const int*)val_s, c->bj_perm, c->bj_pos);
    e = hipGetLastError();
  }
  release();
  if (e != hipSuccess) return fail(c, COUPLB_EDEVICE, hipGetErrorString(e));
  c->bj_npad = npad;
  c->bj_ok = true;
  return COUPLB_OK;
}

int mdf_impl(couplb_ctx* c, int n, const double* x, const double* V, const double* F0,
             double* f_body, const couplb_mdf_args* a, const couplb_mdf_learn* learn,
             couplb_mdf_stats* st) {
  if (n < 0 || !a || (n > 0 && (!x || !V || !f_body))) return fail(c, COUPLB_EINVAL, "bad MDF arguments");
  if (!(a->tol > 0) || a->max_iter < 1 || !(a->relax > 0) ||
      (a->method != COUPLB_MDF_RICHARDSON && a->method != COUPLB_MDF_CG && a->method != COUPLB_MDF_PCG))
    return fail(c, COUPLB_EINVAL, "MDF needs tol > 0, max_iter >= 1, relax > 0 and a known method");
  if (c->nslots != 0 || c->f4_partial)
    return fail(c, COUPLB_ESTATE, "MDF must be the only IBM call between two LB steps");
  if (st) *st = couplb_mdf_stats{0, 0, 0.0, 0.0, -1.0};
  if (n == 0) return COUPLB_OK;
  const bool use_net = learn && learn->warm_start;
  const bool want_x = learn && (learn->warm_start || learn->features || learn->targets);
  if (use_net && learn->warm_start->device != c->p.device)
    return fail(c, COUPLB_EINVAL, "the warm-start network lives on another device");
  const int S = (c->D == 3) ? stencil_size<3>() : stencil_size<2>();
  if ((long)n * S >= (long)INT32_MAX) return fail(c, COUPLB_EINVAL, "too many markers for MDF");
  int rc = ensure_ibm(c, n);
  if (rc) return rc;
  rc = ensure_slot(c, 0);
  if (rc) return rc;
  const int tb = 256, nb = (n + tb - 1) / tb;
  const int nbi = (c->D == 3) ? (4 * n + tb - 1) / tb : nb;
  if (!c->ut) HIPCHK(c, hipMalloc(&c->ut, sizeof(double) * 3 * c->N));
  if (!c->mdf_S) HIPCHK(c, hipMalloc(&c->mdf_S, 2 * sizeof(MdfScalars)));  // CG passes alternate
  if (!c->mdf_host) HIPCHK(c, hipHostMalloc((void**)&c->mdf_host, 2 * sizeof(MdfScalars)));
  for (auto& e : c->mdf_ev)
    if (!e) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  if (c->mdf_cap < n) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->mdf_F); (void)hipFree(c->mdf_R); (void)hipFree(c->mdf_Ap); (void)hipFree(c->mdf_part);
    (void)hipFree(c->mdf_Z);
    c->mdf_Z = nullptr;
    (void)hipFree(c->mm_gam); (void)hipFree(c->mm_B);
    c->mdf_F = c->mdf_R = c->mdf_Ap = c->mdf_part = c->mm_gam = c->mm_B = nullptr;
    c->mdf_cap = 0;
    const int cap = std::max(n, c->cap);
    const int nbmax = (MM_LANES * cap + tb - 1) / tb;  // the widest reduction: k_mm_spmv
    HIPCHK(c, hipMalloc(&c->mdf_F, sizeof(double) * 3 * cap));
    HIPCHK(c, hipMalloc(&c->mdf_R, sizeof(double) * 3 * cap));
    HIPCHK(c, hipMalloc(&c->mdf_Ap, sizeof(double) * 3 * cap));
    HIPCHK(c, hipMalloc(&c->mdf_Z, sizeof(double) * 3 * cap));
    HIPCHK(c, hipMalloc(&c->mdf_part, sizeof(double) * 4 * nbmax));
    HIPCHK(c, hipMalloc(&c->mm_gam, sizeof(double) * cap));
    HIPCHK(c, hipMalloc(&c->mm_B, sizeof(double) * 3 * cap));
    (void)hipFree(c->mlp_x); (void)hipFree(c->mlp_y); (void)hipFree(c->mlp_sig);
    c->mlp_x = c->mlp_y = nullptr;
    c->mlp_sig = nullptr;
    HIPCHK(c, hipMalloc(&c->mlp_x, sizeof(float) * (size_t)MLP_IN * cap));
    HIPCHK(c, hipMalloc(&c->mlp_y, sizeof(float) * 3 * (size_t)cap));
    HIPCHK(c, hipMalloc(&c->mlp_sig, sizeof(double) * (size_t)cap));
    c->mdf_cap = cap;
  }
  double* part_a = c->mdf_part;
  const int nbmax = (MM_LANES * c->mdf_cap + tb - 1) / tb;
  double* part_b = c->mdf_part + nbmax;
  double* part_x = c->mdf_part + 2 * nbmax;  // operator partials inside a CG pass
  double* part_y = c->mdf_part + 3 * nbmax;
  const IbmParams ip = make_ibm_params(c, &a->ibm);
  int tslot = 0;
  rc = timer_begin(c, 3, &tslot);
  if (rc) return rc;
  bool reorder = false;
  rc = prepare_markers(c, n, x, ip, a->ibm.positions_unchanged, false, &reorder);
  if (rc) return rc;
  SlotHost& sl = c->slots[0];
  MdfScalars* Sd = c->mdf_S;
  hipLaunchKernelGGL(k_mdf_reset, dim3(1), dim3(1), 0, c->stream, Sd, 1);
  const double* F0i = want_x ? nullptr : F0;  // the network path starts from F = 0
  if (c->D == 3)
    hipLaunchKernelGGL(k_mdf_init<3>, dim3(nb), dim3(tb), 0, c->stream, n, c->perm, x, V, F0i, ip, c->geo,
                       c->walls, sl.rec, c->mdf_v, c->mdf_F, c->fsrc, c->mm_gam);
  else
    hipLaunchKernelGGL(k_mdf_init<2>, dim3(nb), dim3(tb), 0, c->stream, n, c->perm, x, V, F0i, ip, c->geo,
                       c->walls, sl.rec, c->mdf_v, c->mdf_F, c->fsrc, c->mm_gam);
  HIPCHK(c, hipGetLastError());
  sl.n = n;
  sl.epoch = c->pos_epoch;
  if (c->pair_epoch != sl.epoch) {
    rc = build_pairs(c, sl);
    if (rc) return rc;
  }
  const unsigned nbp = (unsigned)std::max(1, (c->pnh + 255) / 256);
  // The explicit operator (mdf_matrix.hpp) when the pair map is reused from
  // the previous solve (fixed markers): its build is a sort of all triples,
  // paid back over the solves of later steps; the preconditioned solve
  // builds it at once (it saves more passes than the build costs).
  // COUPLB_MDF_MATRIX=0 keeps the node path (and plain CG), =1 builds it at
  // the first solve of a pair map.
  const char* mm_env = getenv("COUPLB_MDF_MATRIX");
  const int mm_mode = mm_env ? atoi(mm_env) : -1;
  const bool pcg_req = a->method == COUPLB_MDF_PCG;
  bool use_mm = false;
  if (mm_mode != 0 && c->pnh > 0 && (mm_mode == 1 || pcg_req || c->mm_seen == c->pair_epoch)) {
    if (c->mm_epoch != c->pair_epoch || c->mm_n != n) {
      rc = mm_build(c, n);
      if (rc) return rc;
    }
    use_mm = c->mm_ok;
  }
  // block-Jacobi preconditioning needs the explicit operator's entries
  if (pcg_req && use_mm && !c->bj_ok) {
    rc = bj_build(c, n, sl.rec);
    if (rc) return rc;
  }
  const bool pcg = pcg_req && use_mm && c->bj_ok;
  c->mm_seen = c->pair_epoch;  // pairs rebuilt for one ordering are the same pairs
  const int nbs = (MM_LANES * n + MDF_THREADS - 1) / MDF_THREADS;
  const int nbr = use_mm ? nbs : nbi;  // partials of a residual / p.Ap pass
  // node values ut = [u*] + S(P)/(2 rho), then the interpolation pass:
  // out = V - I(ut) (affine) or I0(ut); P null stands for zero forces
  auto node_apply = [&](int affine, const double* P, double* out, const MdfScalars* Sx, double* pa,
                        double* pb) {
    hipLaunchKernelGGL(k_mdf_ut, dim3(nbp), dim3(256), 0, c->stream, c->pnh, (const int*)c->pheads,
                       (const int*)c->phlen, (const unsigned*)c->phnode, (const int*)c->pp,
                       (const double*)c->pw, P, (const double*)c->u4, affine, c->ut, Sx);
    if (c->D == 3)
      hipLaunchKernelGGL(k_mdf_interp<3>, dim3(nbi), dim3(MDF_THREADS), 0, c->stream, n, c->geo, c->walls,
                         (const double*)c->ut, affine, (const MarkerRec*)sl.rec, (const double*)c->mdf_v,
                         (const double*)c->fsrc, out, pa, pb, Sx);
    else
      hipLaunchKernelGGL(k_mdf_interp<2>, dim3(nbi), dim3(MDF_THREADS), 0, c->stream, n, c->geo, c->walls,
                         (const double*)c->ut, affine, (const MarkerRec*)sl.rec, (const double*)c->mdf_v,
                         (const double*)c->fsrc, out, pa, pb, Sx);
  };
  // residual R = b - A F (affine, F = rec.f = fsrc on the node path, Fm on the
  // matrix path) or Ap = A P (P = fsrc), with the block partials in pa, pb
  auto apply = [&](int affine, const MdfScalars* Sx, double* pa, double* pb) {
    if (use_mm)
      hipLaunchKernelGGL(k_mm_spmv, dim3(nbs), dim3(MDF_THREADS), 0, c->stream, n,
                         (const int*)c->mm_rowptr, (const int*)c->mm_col, (const double*)c->mm_aval,
                         (const double*)(affine ? c->mdf_F : c->fsrc), (const double*)c->mm_B,
                         affine ? c->mdf_R : c->mdf_Ap, affine, pa, pb, Sx);
    else
      node_apply(affine, c->fsrc, affine ? c->mdf_R : c->mdf_Ap, Sx, pa, pb);
  };
  if (use_mm) {
    // entries for this density, and b = V - I(u*) on the node path
    if (c->mm_q_cap < c->pne) {
      HIPCHK(c, hipStreamSynchronize(c->stream));
      (void)hipFree(c->mm_q);
      c->mm_q = nullptr;
      c->mm_q_cap = 0;
      HIPCHK(c, hipMalloc(&c->mm_q, sizeof(double) * (size_t)c->pne));
      c->mm_q_cap = c->pne;
    }
    hipLaunchKernelGGL(k_mm_q, dim3((unsigned)((c->pne + 255) / 256)), dim3(256), 0, c->stream, c->pne,
                       (const double*)c->pw, (const unsigned*)c->pkey_s, (unsigned)c->N, (const double*)c->u4,
                       c->mm_q);
    hipLaunchKernelGGL(k_mm_values, dim3((unsigned)(((long)MV_LANES * c->mm_nnz + 255) / 256)), dim3(256), 0,
                       c->stream,
                       c->mm_nnz, (const int*)c->mm_ent, (const int*)c->mm_elen, (const int*)c->mm_erow,
                       (const unsigned long long*)c->mm_val, (const double*)c->pw, (const double*)c->mm_q,
                       (const double*)c->mm_gam, c->mm_aval);
    node_apply(1, nullptr, c->mm_B, Sd, part_a, part_b);
    if (pcg)
      hipLaunchKernelGGL(k_bj_invert, dim3((unsigned)(c->bj_npad / BJ)), dim3(64), 0, c->stream,
                         (const int*)c->bj_perm, (const int*)c->bj_pos, (const int*)c->mm_rowptr,
                         (const int*)c->mm_col, (const double*)c->mm_aval, (const double*)c->mm_gam,
                         c->bj_inv);
    HIPCHK(c, hipGetLastError());
  }
  auto scalar = [&](int nblk, int first, int respect, MdfScalars* mirror) {
    hipLaunchKernelGGL(k_mdf_scalar, dim3(1), dim3(MDF_THREADS), 0, c->stream, (const double*)part_a,
                       (const double*)part_b, nblk, a->tol, first, a->max_iter, respect, Sd, mirror);
  };
  auto read = [&]() -> int {
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));  // k_mdf_scalar wrote the pinned mirror
    return COUPLB_OK;
  };
  // true residual of F (rec.f = F): sets done = 1 when converged
  auto true_residual = [&](int first) -> int {
    hipLaunchKernelGGL(k_mdf_setf, dim3(nb), dim3(tb), 0, c->stream, n, (const double*)c->mdf_F, sl.rec,
                       c->fsrc);
    hipLaunchKernelGGL(k_mdf_reset, dim3(1), dim3(1), 0, c->stream, Sd, 0);
    apply(1, Sd, part_a, part_b);
    scalar(nbr, first, 0, c->mdf_host);
    return read();
  };
  constexpr int CHUNK = 8;  // iterations queued between two host checks
  static_assert(CHUNK % 2 == 0, "a CG chunk ends on the first scalar buffer");
  // Chunks of passes until the device sets done: chunk j + 1 is queued before
  // the host waits for chunk j, so the device does not idle on the check; the
  // passes of a chunk queued after convergence return at once.  enqueue(m)
  // queues one chunk whose last kernel writes the scalars to mirror slot m.
  // Ends with the final scalars in slot 0.
  auto run_chunks = [&](auto&& enqueue) -> int {
    MdfScalars* m = c->mdf_host;
    enqueue(m + 0);
    HIPCHK(c, hipEventRecord(c->mdf_ev[0], c->stream));
    for (int j = 0;; j++) {
      enqueue(m + ((j + 1) & 1));
      HIPCHK(c, hipGetLastError());
      HIPCHK(c, hipEventRecord(c->mdf_ev[(j + 1) & 1], c->stream));
      HIPCHK(c, hipEventSynchronize(c->mdf_ev[j & 1]));
      if (m[j & 1].done) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        m[0] = m[j & 1];
        return COUPLB_OK;
      }
      if (j > a->max_iter / CHUNK + 4) {  // every pass counts, so max_iter ends the loop first
        (void)hipStreamSynchronize(c->stream);
        return fail(c, COUPLB_ESTATE, "MDF iteration did not end");
      }
    }
  };
  rc = true_residual(1);
  if (rc) return rc;
  double slip_zero = -1.0;
  if (want_x) {
    slip_zero = c->mdf_host->smax;
    if (c->D == 3)
      hipLaunchKernelGGL(k_mlp_features<3>, dim3(nb), dim3(tb), 0, c->stream, n, c->geo, c->walls,
                         (const MarkerRec*)sl.rec, (const double*)c->u4, (const double*)c->mdf_v,
                         (const double*)c->mdf_R, c->mlp_x, c->mlp_sig);
    else
      hipLaunchKernelGGL(k_mlp_features<2>, dim3(nb), dim3(tb), 0, c->stream, n, c->geo, c->walls,
                         (const MarkerRec*)sl.rec, (const double*)c->u4, (const double*)c->mdf_v,
                         (const double*)c->mdf_R, c->mlp_x, c->mlp_sig);
    HIPCHK(c, hipGetLastError());
    if (learn->features) {
      const long ne = (long)n * MLP_IN;
      hipLaunchKernelGGL(k_mlp_unsort, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c->stream, n,
                         (const int*)c->perm, (const float*)c->mlp_x, learn->features);
    }
    if (use_net && c->mdf_host->done == 0) {
      couplb::mlp::forward(n, c->mlp_x, learn->warm_start->params, c->mlp_y, c->stream);
      hipLaunchKernelGGL(k_mdf_warm, dim3(nb), dim3(tb), 0, c->stream, n, (const float*)c->mlp_y,
                         (const double*)c->mlp_sig, c->mdf_F, sl.rec, c->fsrc);
      HIPCHK(c, hipGetLastError());
      rc = true_residual(1);  // slip at the prediction: where the iteration starts
      if (rc) return rc;
    }
  }
  int method = a->method;
  while (c->mdf_host->done == 0) {
    if (method != COUPLB_MDF_RICHARDSON) {
      if (pcg) {  // z = M^-1 r, p = z, rz = r.z
        hipLaunchKernelGGL((k_pcg_step<false>), dim3(nb), dim3(MDF_THREADS), 0, c->stream, c->bj_npad,
                           (const int*)c->bj_perm, (const double*)c->bj_inv, c->fsrc, (const double*)c->mdf_Ap,
                           c->mdf_F, c->mdf_R, c->mdf_Z, (const double*)part_x, 0, part_a, part_b, Sd);
        hipLaunchKernelGGL(k_pcg_rz, dim3(1), dim3(MDF_THREADS), 0, c->stream, (const double*)part_a, nb, Sd);
      } else {
        hipLaunchKernelGGL(k_cg_restart, dim3((unsigned)((3 * n + tb - 1) / tb)), dim3(tb), 0, c->stream, n,
                           (const double*)c->mdf_R, c->fsrc);
      }
      // pass k reads the scalars Sd[k % 2] and writes Sd[(k + 1) % 2]
      rc = run_chunks([&](MdfScalars* mirror) {
        for (int k = 0; k < CHUNK; k++) {
          MdfScalars* Sk = Sd + (k & 1);
          MdfScalars* Sn = Sd + ((k + 1) & 1);
          apply(0, Sk, part_x, part_y);
          if (pcg)
            hipLaunchKernelGGL((k_pcg_step<true>), dim3(nb), dim3(MDF_THREADS), 0, c->stream, c->bj_npad,
                               (const int*)c->bj_perm, (const double*)c->bj_inv, c->fsrc,
                               (const double*)c->mdf_Ap, c->mdf_F, c->mdf_R, c->mdf_Z, (const double*)part_x,
                               nbr, part_a, part_b, Sk);
          else
            hipLaunchKernelGGL(k_cg_update, dim3(nb), dim3(MDF_THREADS), 0, c->stream, n,
                               (const double*)c->fsrc, (const double*)c->mdf_Ap, c->mdf_F, c->mdf_R,
                               (const double*)part_x, nbr, part_a, part_b, Sk);
          // p = r + beta p (CG) or z + beta p (preconditioned)
          hipLaunchKernelGGL(k_cg_next, dim3(nb), dim3(MDF_THREADS), 0, c->stream, n,
                             (const double*)(pcg ? c->mdf_Z : c->mdf_R),
                             c->fsrc, (const double*)part_a, (const double*)part_b, nb, a->tol, a->max_iter,
                             (const MdfScalars*)Sk, Sn, k == CHUNK - 1 ? mirror : nullptr);
        }
      });
      if (rc) return rc;
      const int why = c->mdf_host->done;
      if (c->mdf_host->iters >= a->max_iter) {
        rc = true_residual(0);
        if (rc) return rc;
        break;
      }
      rc = true_residual(0);  // recursive residual converged or CG broke down
      if (rc) return rc;
      if (c->mdf_host->done == 0 && why == 2) method = COUPLB_MDF_RICHARDSON;
    } else {
      rc = run_chunks([&](MdfScalars* mirror) {
        for (int k = 0; k < CHUNK; k++) {
          hipLaunchKernelGGL(k_mdf_relax, dim3(nb), dim3(tb), 0, c->stream, n, (const double*)c->mdf_R,
                             a->relax, c->mdf_F, (const MarkerRec*)sl.rec, c->fsrc, (const MdfScalars*)Sd);
          apply(1, Sd, part_a, part_b);
          scalar(nbr, 0, 1, k == CHUNK - 1 ? mirror : nullptr);
        }
      });
      if (rc) return rc;
    }
  }
  const MdfScalars& hs = *c->mdf_host;
  couplb_mdf_stats out{hs.iters, hs.done == 1 ? 1 : 0, hs.smax_first, hs.smax, slip_zero};
  if (learn && learn->targets)
    hipLaunchKernelGGL(k_mdf_targets, dim3(nb), dim3(tb), 0, c->stream, n, (const int*)c->perm,
                       (const double*)c->mdf_F, (const double*)c->mlp_sig, (const MarkerRec*)sl.rec,
                       learn->targets);
  hipLaunchKernelGGL(k_mdf_setf, dim3(nb), dim3(tb), 0, c->stream, n, (const double*)c->mdf_F, sl.rec,
                     c->fsrc);
  hipLaunchKernelGGL(k_mdf_out, dim3(nb), dim3(tb), 0, c->stream, n, c->perm, (const double*)c->mdf_F,
                     (const MarkerRec*)sl.rec, a->ibm.force_scale, f_body);
  HIPCHK(c, hipGetLastError());
  rc = timer_end(c, tslot);
  if (rc) return rc;
  c->nslots = 1;
  c->rho_fresh = 1;  // the velocity gather is compute_macroscopic(false) of this state
  if (st) *st = out;
  return COUPLB_OK;
}

// The pending record slots spread into F4 / flag (spread_rows.hpp): the
// records of all slots binned by base row and i0 segment (counting sort,
// each bucket ordered by slot and marker), then one wavefront per tile of
// nodes.  fold: add the periodic ghost images into the interior; accumulate:
// add to the forces F4 / flag hold.  moved_only: only the records that left
// their ordered base cell (the moving-marker map holds the others), only the
// tiles they reach, F4 and flag written only where they add a force.
// bin: when given, launches pass 1 itself (spread_rows.hpp bin_record for
// every record) in place of k_rec_count.
int row_spread(couplb_ctx* c, double* F4, unsigned char* flag, int fold, int accumulate,
               bool moved_only = false,
               const std::function<int(const SlotSpan&, const RecBin&)>* bin = nullptr) {
  SlotSpan ss;
  ss.bases = moved_only ? (const Base*)c->bases : nullptr;
  ss.nslots = c->nslots;
  ss.off[0] = 0;
  for (int s = 0; s < c->nslots; s++) {
    ss.rec[s] = c->slots[s].rec;
    ss.off[s + 1] = ss.off[s] + std::max(0, c->slots[s].n);
  }
  for (int s = c->nslots; s < MAX_SLOTS; s++) ss.rec[s] = nullptr;
  const long total = ss.off[c->nslots];
  const int nb = c->geo.nrows * spread_nxb(c->geo);
  if (nb != c->sb_nb) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->sb_cnt); (void)hipFree(c->sb_start); (void)hipFree(c->sb_scan_tmp);
    c->sb_cnt = c->sb_start = nullptr;
    c->sb_scan_tmp = nullptr;
    c->sb_nb = 0;
    HIPCHK(c, hipMalloc(&c->sb_cnt, sizeof(int) * ((size_t)nb + 1)));
    HIPCHK(c, hipMalloc(&c->sb_start, sizeof(int) * ((size_t)nb + 1)));
    size_t bytes = 0;
    HIPCHK(c, rocprim::exclusive_scan(nullptr, bytes, c->sb_cnt, c->sb_start, 0, (size_t)nb + 1,
                                      rocprim::plus<int>(), c->stream));
    HIPCHK(c, hipMalloc(&c->sb_scan_tmp, bytes));
    c->sb_scan_bytes = bytes;
    c->sb_nb = nb;
  }
  if (total > c->sb_cap || c->sb_cap == 0) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->sb_bk); (void)hipFree(c->sb_rank); (void)hipFree(c->sb_ent); (void)hipFree(c->sb_ent2);
    (void)hipFree(c->sb_large);
    c->sb_bk = c->sb_rank = c->sb_large = nullptr;
    c->sb_ent = c->sb_ent2 = nullptr;
    const long cap = std::max(std::max(total, 2 * c->sb_cap), 64L);
    c->sb_cap = 0;
    HIPCHK(c, hipMalloc(&c->sb_bk, sizeof(int) * cap));
    HIPCHK(c, hipMalloc(&c->sb_rank, sizeof(int) * cap));
    HIPCHK(c, hipMalloc(&c->sb_ent, sizeof(unsigned long long) * cap));
    HIPCHK(c, hipMalloc(&c->sb_ent2, sizeof(unsigned long long) * cap));
    // at most cap / (SB_SMALL + 1) buckets are large; their count follows
    HIPCHK(c, hipMalloc(&c->sb_large, sizeof(int) * ((size_t)cap / (SB_SMALL + 1) + 2)));
    c->sb_cap = cap;
  }
  const int tb = 256;
  int* nlarge = c->sb_large + ((size_t)c->sb_cap / (SB_SMALL + 1) + 1);
  hipLaunchKernelGGL(k_fill_int, dim3((unsigned)(((size_t)nb + 1 + tb - 1) / tb)), dim3(tb), 0, c->stream, c->sb_cnt,
                     (size_t)nb + 1, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, nlarge, (size_t)1, 0);
  HIPCHK(c, hipGetLastError());
  // tiles of the spreading (TileShape) and, for moved_only, their marks
  int ntx, nty, ntz;
  if (c->D == 3) {
    typedef TileShape<3> T;
    ntx = (c->geo.nx + T::X - 1) / T::X; nty = (c->geo.ny + T::Y - 1) / T::Y; ntz = (c->geo.nz + T::Z - 1) / T::Z;
  } else {
    typedef TileShape<2> T;
    ntx = (c->geo.nx + T::X - 1) / T::X; nty = (c->geo.ny + T::Y - 1) / T::Y; ntz = 1;
  }
  const int ntiles = ntx * nty * ntz;
  if (moved_only && ntiles != c->ntiles) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(c->tflag); (void)hipFree(c->tlist);
    c->tflag = nullptr; c->tlist = nullptr; c->ntiles = 0;
    HIPCHK(c, hipMalloc(&c->tflag, (size_t)ntiles));
    HIPCHK(c, hipMalloc(&c->tlist, sizeof(int) * ((size_t)ntiles + 1)));
    hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks((size_t)ntiles)), dim3(256), 0, c->stream, c->tflag,
                       (size_t)ntiles, (unsigned char)0);
    c->ntiles = ntiles;
  }
  if (moved_only)
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, c->stream, c->tlist + ntiles, (size_t)1, 0);
  const RecBin rb{c->sb_cnt, c->sb_bk, c->sb_rank, moved_only ? c->tflag : nullptr, fold};
  if (bin) {
    const int rc = (*bin)(ss, rb);
    if (rc) return rc;
  } else if (total > 0) {
    const unsigned gb = (unsigned)((total + tb - 1) / tb);
    if (c->D == 3)
      hipLaunchKernelGGL(k_rec_count<3>, dim3(gb), dim3(tb), 0, c->stream, ss, c->geo, rb);
    else
      hipLaunchKernelGGL(k_rec_count<2>, dim3(gb), dim3(tb), 0, c->stream, ss, c->geo, rb);
    HIPCHK(c, hipGetLastError());
  }
  size_t bytes = c->sb_scan_bytes;
  HIPCHK(c, rocprim::exclusive_scan(c->sb_scan_tmp, bytes, c->sb_cnt, c->sb_start, 0, (size_t)nb + 1,
                                    rocprim::plus<int>(), c->stream));
  if (total > 0) {
    const unsigned gb = (unsigned)((total + tb - 1) / tb);
    if (c->D == 3)
      hipLaunchKernelGGL(k_rec_place<3>, dim3(gb), dim3(tb), 0, c->stream, ss, c->geo, c->walls, fold,
                         (const int*)c->sb_bk, (const int*)c->sb_rank, (const int*)c->sb_start, c->sb_ent,
                         rb.tflag);
    else
      hipLaunchKernelGGL(k_rec_place<2>, dim3(gb), dim3(tb), 0, c->stream, ss, c->geo, c->walls, fold,
                         (const int*)c->sb_bk, (const int*)c->sb_rank, (const int*)c->sb_start, c->sb_ent,
                         rb.tflag);
    hipLaunchKernelGGL(k_bucket_sort, dim3((unsigned)((nb + tb - 1) / tb)), dim3(tb), 0, c->stream, nb,
                       (const int*)c->sb_start, c->sb_ent, c->sb_large, nlarge);
    hipLaunchKernelGGL(k_bucket_sort_large, dim3(256), dim3(tb), 0, c->stream, (const int*)c->sb_large,
                       (const int*)nlarge, (const int*)c->sb_start, c->sb_ent, c->sb_ent2);
    HIPCHK(c, hipGetLastError());
  }
  if (moved_only) {
    // the number of records off the map goes to the host for the decision
    // to renew the ordering (ibm_impl); no wait for it
    if (!c->moved_host) {
      HIPCHK(c, hipHostMalloc((void**)&c->moved_host, sizeof(int)));
      HIPCHK(c, hipEventCreateWithFlags(&c->moved_ev, hipEventDisableTiming));
    }
    const bool report = !c->moved_pending;
    hipLaunchKernelGGL(k_tile_list, dim3((unsigned)((ntiles + tb - 1) / tb)), dim3(tb), 0, c->stream, ntiles,
                       c->tflag, c->tlist, c->tlist + ntiles, (const int*)(c->sb_start + nb),
                       report ? c->moved_host : nullptr);
    if (report) {
      HIPCHK(c, hipEventRecord(c->moved_ev, c->stream));
      c->moved_pending = 1;
      c->moved_epoch = c->pos_epoch;
      c->moved_total = total;
    }
  }
  TileSpreadArgs ta;
  ta.tag = (unsigned char)(moved_only ? c->f2_tag : 1);
  ta.tiles = moved_only ? c->tlist : nullptr;
  ta.ntiles = moved_only ? c->tlist + ntiles : nullptr;
  ta.g = c->geo;
  ta.w = c->walls;
  ta.bstart = c->sb_start;
  ta.ent = c->sb_ent;
  ta.ss = ss;
  ta.F4 = F4;
  ta.flag = flag;
  ta.fold = fold;
  ta.accumulate = accumulate;
  // a periodic axis of fewer than 4 nodes lets one stencil reach a node and
  // its ghost image together
  ta.serial = 0;
  const int ext[3] = {c->geo.nx, c->geo.ny, c->geo.nz};
  for (int d = 0; d < c->D; d++)
    if (fold && ext[d] < 4 && (c->walls.type[d][0] == 0 || c->walls.type[d][1] == 0)) ta.serial = 1;
  // moved_only: a fixed grid walks the tile list
  const dim3 grid = moved_only ? dim3((unsigned)std::min(ntiles, 2048), 1u, 1u)
                               : dim3((unsigned)ntx, (unsigned)nty, (unsigned)ntz);
  if (c->D == 3)
    hipLaunchKernelGGL(k_spread_tiles<3>, grid, dim3(64), 0, c->stream, ta);
  else
    hipLaunchKernelGGL(k_spread_tiles<2>, grid, dim3(64), 0, c->stream, ta);
  HIPCHK(c, hipGetLastError());
  return COUPLB_OK;
}

// All record slots in use: spread them into F4 (added to what it holds)
// and free the slots for the following sub-steps of the interval.  Any
// number of sub-steps per LB step is then spread, in sub-step order.
int fold_slots(couplb_ctx* c) {
  const int rc = row_spread(c, c->F4, c->flag, 1, c->f4_partial);
  if (rc) return rc;
  c->pair_epoch = -1;  // the node flags now describe F4
  c->mpair_epoch = -1;
  c->f4_partial = 1;
  c->nslots = 0;
  return COUPLB_OK;
}

// Arguments of the next LB step of the current state.
StepArgs step_args(couplb_ctx* c, const double g[3], int flags) {
  StepArgs a;
  a.pc = make_pull(c);
  a.dst = c->f[c->cur ^ 1];
  a.omega = 1.0 / c->p.tau;
  const double ZERO_TOL = 1e-15;  // fix_couplb.cpp:693-695
  const bool grav = g && !(std::fabs(g[0]) < ZERO_TOL && std::fabs(g[1]) < ZERO_TOL &&
                           std::fabs(g[2]) < ZERO_TOL);
  for (int d = 0; d < 3; d++) a.g[d] = grav ? g[d] : 0.0;
  a.grav_mode = grav ? (c->rho_fresh ? 1 : 2) : 0;
  a.write_rho = (grav && !(flags & COUPLB_STEP_RHO_REFRESHED_NEXT)) ? 1 : 0;
  a.has_moving = c->has_moving;
  a.rho_out = c->rho;
  a.has_ibm = c->nslots > 0 || c->f4_partial;
  a.flag = c->flag;
  a.F4 = c->F4;
  a.pair_mode = 0;
  a.plen = nullptr;
  a.pseg = nullptr;
  a.pp = nullptr;
  a.pw = nullptr;
  a.fsrc = nullptr;
  a.pcon = nullptr;
  a.pidx = nullptr;
  a.flag2 = nullptr;
  a.has_f2 = 0;
  return a;
}

// Plain copy in 16-byte groups, four groups in flight per lane, one pass
// over the grid: the achievable-bandwidth reference for the kernels'
// roofline fractions.
typedef unsigned int CopyV __attribute__((ext_vector_type(4)));
constexpr int COPY_UNROLL = 4;
__global__ void __launch_bounds__(256) k_copy16(const CopyV* __restrict__ src, CopyV* __restrict__ dst, size_t n) {
  const size_t base = (size_t)blockIdx.x * (256 * COPY_UNROLL) + threadIdx.x;
  CopyV v[COPY_UNROLL];
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; u++)
    if (base + (size_t)u * 256 < n) v[u] = __builtin_nontemporal_load(src + base + (size_t)u * 256);
#pragma unroll
  for (int u = 0; u < COPY_UNROLL; u++)
    if (base + (size_t)u * 256 < n) __builtin_nontemporal_store(v[u], dst + base + (size_t)u * 256);
}

template <int D, typename S> struct LaunchFused {
  static void run(couplb_ctx* c, const StepArgs& a, int n) {
    const unsigned t = (unsigned)std::min<size_t>(FUSED_THREADS, round_up(c->N, 64));
    hipLaunchKernelGGL((k_steps_fused<D, S>), dim3(1), dim3(t), c->fused_lds, c->stream, a, n);
  }
};

// LDS bytes of k_steps_fused for this grid, 0 when the grid is larger than
// the node bound or the state does not fit the LDS of one workgroup
template <int D, typename S> struct FusedSetup {
  static void run(couplb_ctx* c, size_t* bytes) {
    *bytes = 0;
    if (c->N > c->fused_nodes || c->N > FUSED_THREADS * 4) return;
    const size_t b = fused_lds_bytes<D, S>(c->N);
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    if (b > (size_t)lim) return;
    if (b > 65536 && hipFuncSetAttribute(reinterpret_cast<const void*>(&k_steps_fused<D, S>),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)b) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    *bytes = b;
  }
};

// n steps after a first one (no pending IBM force, state in pull form) as
// k_steps_fused launches of at most FUSED_MAX_RUN steps (a bound on the
// duration of one launch); the context ends where n step_impl calls leave it.
constexpr int FUSED_MAX_RUN = 1 << 16;
int fused_steps(couplb_ctx* c, const double g[3], int flags, int n) {
  if (c->nslots > 0 || c->f4_partial || c->local) return COUPLB_EINVAL;
  while (n > 0) {
    const int m = std::min(n, FUSED_MAX_RUN);
    const StepArgs a = step_args(c, g, flags);
    dispatch<LaunchFused>(c, a, m);
    HIPCHK(c, hipGetLastError());
    c->cur ^= (m & 1);
    c->prev_valid = 1;
    c->rho_fresh = 0;
    c->cycle += m;
    n -= m;
  }
  return COUPLB_OK;
}

int step_impl(couplb_ctx* c, const double g[3], int flags, bool timed) {
  StepArgs a = step_args(c, g, flags);
  const int dst = c->cur ^ 1;
  const int S = (c->D == 3) ? stencil_size<3>() : stencil_size<2>();
  // markers at the positions of the previous LB interval: constant map
  bool fixed = !c->f4_partial && c->nslots == 1 && c->slots[0].epoch == c->prev_step_epoch &&
                     !c->kept_move && (long)c->slots[0].n * S < (long)INT32_MAX;
  // moving markers whose sub-steps all kept one ordering: the moving-marker
  // map of that ordering, the records that left their base cell spread apart
  bool moving = !fixed && !c->f4_partial && c->nslots > 0 && c->order_n > 0 &&
                (long)c->order_n * S < (long)INT32_MAX;
  for (int s = 0; moving && s < c->nslots; s++)
    moving = c->slots[s].epoch == c->pos_epoch && c->slots[s].n == c->order_n;
  int tspread = -1;
  if ((moving || fixed) && timed) {
    const int rc = timer_begin(c, 2, &tspread);
    if (rc) return rc;
  }
  if (moving && c->mpair_epoch != c->pos_epoch) {
    const int rc = build_moving_map(c, c->order_n);
    if (rc) return rc;
  } else if (fixed && c->pair_epoch != c->slots[0].epoch) {
    const int rc = build_pairs(c, c->slots[0]);
    if (rc) return rc;
  }
  // a node with more pairs than a run length holds (some thousand markers
  // per cell): spread through the node rows instead
  if ((moving || fixed) && c->runs_long) moving = fixed = false;
  if (moving) {
    // the pair forces of the records on their ordered base cell (the same
    // kernel bins the others), then those others spread into F4 where flag2
    // is set
    // flag2 marks with a tag that changes every step (1..255), so it is
    // cleared only when the tags come round again
    if (!c->flag2) HIPCHK(c, hipMalloc(&c->flag2, c->N));
    c->f2_tag = (c->f2_tag % 255) + 1;
    if (c->f2_tag == 1)
      hipLaunchKernelGGL(k_fill_u8, dim3(fill_u8_blocks(c->N)), dim3(256), 0, c->stream, c->flag2, c->N,
                         (unsigned char)0);
    const std::function<int(const SlotSpan&, const RecBin&)> con = [c, S](const SlotSpan& ss, const RecBin& rb) {
      const long nt = (long)c->order_n * ((c->D == 3) ? 4 : 1);  // threads: one per stencil plane
      const int tb = 256;
      if (c->D == 3)
        hipLaunchKernelGGL(k_pair_con<3>, dim3((unsigned)((nt + tb - 1) / tb)), dim3(tb), 0, c->stream,
                           c->order_n, ss, c->pcon, c->geo, rb);
      else
        hipLaunchKernelGGL(k_pair_con<2>, dim3((unsigned)((nt + tb - 1) / tb)), dim3(tb), 0, c->stream,
                           c->order_n, ss, c->pcon, c->geo, rb);
      HIPCHK(c, hipGetLastError());
      return COUPLB_OK;
    };
    const int rc = row_spread(c, c->F4, c->flag2, 1, 0, true, &con);
    if (rc) return rc;
    a.pair_mode = 2;
    a.plen = c->plen;
    a.pseg = c->pseg;
    a.pcon = c->pcon;
    a.pidx = c->pval_s;
    a.flag2 = c->flag2;
    a.has_f2 = c->f2_tag;
  } else if (fixed) {
    // the step sums each node's pair run itself (no node-force pass)
    a.pair_mode = 1;
    a.plen = c->plen;
    a.pseg = c->pseg;
    a.pp = c->pp;
    a.pw = c->pw;
    a.fsrc = c->fsrc;
  } else if (c->nslots > 0) {
    // spread the pending marker records into the per-node force
    c->pair_epoch = -1;  // the spreading rewrites the node flags
    c->mpair_epoch = -1;
    if (timed && tspread < 0) {
      const int rc = timer_begin(c, 2, &tspread);
      if (rc) return rc;
    }
    const int rc = row_spread(c, c->F4, c->flag, 1, c->f4_partial);
    if (rc) return rc;
  }
  if (tspread >= 0) {
    const int rc = timer_end(c, tspread);
    if (rc) return rc;
  }
  int tstep = -1;
  if (timed) {
    const int rc = timer_begin(c, 0, &tstep);
    if (rc) return rc;
  }
  dispatch<LaunchStep>(c, a);
  HIPCHK(c, hipGetLastError());
  if (tstep >= 0) {
    const int rc = timer_end(c, tstep);
    if (rc) return rc;
  }
  if (c->nslots > 0) {
    c->prev_step_epoch = (c->nslots == 1) ? c->slots[0].epoch : -1;
    c->order_age++;
  }
  c->cur = dst;
  c->local = 0;
  c->prev_valid = 1;
  c->nslots = 0;
  c->f4_partial = 0;
  c->rho_fresh = 0;
  c->cycle++;
  return COUPLB_OK;
}

// A host-array sub-step that only recomputes forces (force_only): v (and
// m, f where needed) go up, k_marker_force runs over the atoms in caller
// order, and the forces come back.  By default all of it is on the context
// stream: a kernel waiting on another stream's event has started up to
// 0.4 ms after the event here (measured), more than the copies take.
// COUPLB_HOST_CHUNKS = K > 1 instead splits the atoms into K chunks with
// uploads on the copy stream and downloads on the download stream, each
// chunk's kernel waiting for its upload (overlap of the copies, at the
// price of those waits and of 35-40 us of host calls per chunk).
int host_force_chunks(couplb_ctx* c, int n, const double* v, const double* m, double* f,
                      const couplb_ibm_args* a, const IbmParams& ip) {
  static const int env_chunks = [] {
    const char* e = getenv("COUPLB_HOST_CHUNKS");
    return e ? atoi(e) : 1;
  }();
  const int K = n < 8192 ? 1 : std::max(1, std::min(env_chunks, HOST_CHUNKS_MAX));
  if (K > 1 && !c->dstream) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->dstream, hipStreamNonBlocking));
    for (int k = 0; k < HOST_CHUNKS_MAX; k++) {
      HIPCHK(c, hipEventCreateWithFlags(&c->ev_up[k], hipEventDisableTiming));
      HIPCHK(c, hipEventCreateWithFlags(&c->ev_cmp[k], hipEventDisableTiming));
    }
  }
  hipStream_t up = K > 1 ? c->cstream : c->stream, down = K > 1 ? c->dstream : c->stream;
  const int tb = 256;
  if (c->iperm_epoch != c->pos_epoch) {
    hipLaunchKernelGGL(k_inv_perm, dim3((unsigned)((n + tb - 1) / tb)), dim3(tb), 0, c->stream, n,
                       (const int*)c->perm, c->iperm);
    HIPCHK(c, hipGetLastError());
    c->iperm_epoch = c->pos_epoch;
  }
  const bool send_m = !a->masses_unchanged;  // dm holds the previous sub-step's masses
  for (int k = 0; k < K; k++) {
    const size_t lo = (size_t)n * k / K, cnt = (size_t)n * (k + 1) / K - lo;
    HIPCHK(c, hipMemcpyAsync(c->dv + 3 * lo, v + 3 * lo, sizeof(double) * 3 * cnt, hipMemcpyHostToDevice, up));
    if (send_m)
      HIPCHK(c, hipMemcpyAsync(c->dm + lo, m + lo, sizeof(double) * cnt, hipMemcpyHostToDevice, up));
    if (!a->forces_zero)
      HIPCHK(c, hipMemcpyAsync(c->df + 3 * lo, f + 3 * lo, sizeof(double) * 3 * cnt, hipMemcpyHostToDevice,
                               up));
    if (K > 1) HIPCHK(c, hipEventRecord(c->ev_up[k], up));
  }
  c->hx_n = -1;
  c->host_cycle = c->cycle;
  int tslot = 0;
  int rc = timer_begin(c, 1, &tslot);
  if (rc) return rc;
  SlotHost& sl = c->slots[c->nslots - 1];
  double* fsrc = c->nslots == 1 ? c->fsrc : nullptr;
  for (int k = 0; k < K; k++) {
    const size_t lo = (size_t)n * k / K, cnt = (size_t)n * (k + 1) / K - lo;
    if (K > 1) HIPCHK(c, hipStreamWaitEvent(c->stream, c->ev_up[k], 0));
    hipLaunchKernelGGL(k_marker_force, dim3((unsigned)((cnt + tb - 1) / tb)), dim3(tb), 0, c->stream, (int)cnt,
                       (const int*)c->perm, (const double*)c->dv, (const double*)c->dm, c->df,
                       a->forces_zero ? 1 : 0, ip, (const double*)c->ufc, sl.rec, fsrc, (int)lo,
                       (const int*)c->iperm);
    HIPCHK(c, hipGetLastError());
    if (K > 1) {
      HIPCHK(c, hipEventRecord(c->ev_cmp[k], c->stream));
      HIPCHK(c, hipStreamWaitEvent(down, c->ev_cmp[k], 0));
    }
    HIPCHK(c, hipMemcpyAsync(f + 3 * lo, c->df + 3 * lo, sizeof(double) * 3 * cnt, hipMemcpyDeviceToHost,
                             down));
  }
  rc = timer_end(c, tslot);
  if (rc) return rc;
  c->rho_fresh = 1;  // compute_macroscopic(false) at fix_couplb.cpp:758
  HIPCHK(c, hipStreamSynchronize(down));
  c->hx_n = n;
  return COUPLB_OK;
}

}  // namespace

extern "C" {

const char* couplb_version(void) { return "couplb-mi355x 0.1 (gfx950)"; }

const char* couplb_last_error(const couplb_ctx* c) { return c ? c->err.c_str() : "null context"; }

int couplb_create(const couplb_params* p, couplb_ctx** out) {
  if (!p || !out) return COUPLB_EINVAL;
  *out = nullptr;
  auto* c = new couplb_ctx();
  c->p = *p;
  auto bad = [&](const char* m) { c->err = m; *out = c; return COUPLB_EINVAL; };
  if (p->dim != 2 && p->dim != 3) return bad("dim must be 2 or 3");
  const int nz = p->dim == 3 ? p->nz : 1;
  if (p->nx < 2 || p->ny < 2 || (p->dim == 3 && p->nz < 2) || (p->dim == 2 && p->nz != 1))
    return bad("grid extents must be >= 2 (nz = 1 in 2D)");
  if (!(p->tau > 0.5)) return bad("tau <= 0.5 (unstable)");
  if (p->kernel != COUPLB_ROMA && p->kernel != COUPLB_PESKIN4) return bad("unknown IBM kernel");
  if (p->storage < COUPLB_FP64 || p->storage > COUPLB_I16) return bad("unknown storage");
  if (p->storage == COUPLB_I16 && !(p->storage_range >= 0.0 && p->storage_range < 1e6))
    return bad("storage_range must be >= 0 (0: default)");
  if (!(p->rho0 > 0)) return bad("rho0 must be positive");
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++)
      if (p->wall[d][s] < 0 || p->wall[d][s] > 4) return bad("wall types must be in [0,4]");
  if (p->dim == 2 && (p->wall[2][0] || p->wall[2][1])) return bad("2D grid with z walls");
  if ((size_t)p->nx * p->ny * nz >= (1ull << 32)) return bad("grid too large for 32-bit node index");
  c->D = p->dim;
  c->Q = p->dim == 3 ? 19 : 9;
  c->geo.nx = p->nx; c->geo.ny = p->ny; c->geo.nz = nz; c->geo.D = p->dim;
  c->geo.nry = p->ny + 5;
  c->geo.nrows = c->geo.nry * (p->dim == 3 ? nz + 5 : 1);
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      c->walls.type[d][s] = (d < p->dim) ? p->wall[d][s] : 0;
      for (int k = 0; k < 3; k++) c->walls.vel[d][s][k] = p->face_vel[d][s][k];
      if (d < p->dim && p->wall[d][s] == 2) c->has_moving = 1;
    }
  c->N = (size_t)p->nx * p->ny * nz;
  // Population layout.  rows (default): each x-row keeps its Q segments
  // together (segment nx rounded up to 8, COUPLB_ROW_PAD elements after each
  // row, 256 by default), so a block's loads and stores stay within five
  // neighbouring rows: 3 % faster than planes for the FP64 step.  planes:
  // one q plane after the other, each whole 4 KiB pages plus a skew so the
  // plane streams do not start on the same HBM channel (COUPLB_PLANE_PAD
  // overrides the skew).  COUPLB_LAYOUT=planes selects the latter.
  {
    const char* e = std::getenv("COUPLB_LAYOUT");
    const bool rows = e ? std::strcmp(e, "planes") != 0 : true;
    const size_t nrows = (size_t)p->ny * nz;
    if (rows) {
      size_t pad = 256;
      if (const char* v = std::getenv("COUPLB_ROW_PAD")) pad = (size_t)std::strtoull(v, nullptr, 10);
      c->lay.pq = round_up((size_t)p->nx, 8);
      c->lay.pr = c->lay.pq * (size_t)c->Q + pad;
      c->falloc = c->lay.pr * nrows;
    } else {
      size_t pad = 256;
      if (const char* v = std::getenv("COUPLB_PLANE_PAD")) pad = (size_t)std::strtoull(v, nullptr, 10);
      c->lay.pq = round_up(c->N, 512) + pad;
      c->lay.pr = (size_t)p->nx;
      c->falloc = c->lay.pq * (size_t)c->Q;
    }
  }
  // several steps of a grid this small run as one launch (k_steps_fused);
  // COUPLB_FUSED_NODES overrides the bound (0 switches the fusion off)
  c->fused_nodes = 1024;
  if (const char* v = std::getenv("COUPLB_FUSED_NODES")) c->fused_nodes = (size_t)std::strtoull(v, nullptr, 10);
  // LB intervals a moving-marker ordering is kept (0: a new ordering every
  // interval)
  if (const char* v = std::getenv("COUPLB_REORDER_EVERY")) c->reorder_every = (int)std::strtol(v, nullptr, 10);
  c->elem = storage_bytes(p->storage);
  c->scale = storage_scale(p);
  c->base = (p->storage == COUPLB_FP64) ? 0.0 : p->rho0;
  int dev = p->device;
  if (hipSetDevice(dev) != hipSuccess) return bad("hipSetDevice failed");
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
    return bad("hipStreamCreate failed");
  c->scratch = scratch_on(dev);
  if (!c->scratch) return bad("hipGetSymbolAddress of the diagnostics scratch failed");
  for (int e = 0; e < TimerRing::R; e++)
    if (hipEventCreate(&c->tr.a[e]) != hipSuccess || hipEventCreate(&c->tr.b[e]) != hipSuccess)
      return bad("hipEventCreate failed");
  dispatch<FusedSetup>(c, &c->fused_lds);
  const size_t fbytes = c->elem * c->falloc;
  for (int b = 0; b < 2; b++) {
    if (hipMalloc(&c->f[b], fbytes) != hipSuccess) {
      c->err = "hipMalloc of populations failed";
      *out = c;
      return COUPLB_ENOMEM;
    }
  }
  if (hipMalloc(&c->rho, sizeof(double) * c->N) != hipSuccess) {
    c->err = "hipMalloc of density failed";
    *out = c;
    return COUPLB_ENOMEM;
  }
  // init_equilibrium(rho_lb, 0, 0, 0) (couplb_lattice.h:145-154): f = w_q
  // rho0 (zero deviation); density array = rho0 (fix_couplb.cpp:504).
  {
    const int gx = p->nx + 2, gy = p->ny + 2, gz = p->dim == 3 ? nz + 2 : 1;
    const size_t nt = (size_t)gx * gy * gz;
    std::vector<double> h((size_t)c->Q * nt);
    const double* w3 = nullptr;
    static const double W3[19] = {1.0/3.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0, 1.0/18.0,
                                  1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0,
                                  1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0};
    static const double W2[9] = {4.0/9.0, 1.0/9.0, 1.0/9.0, 1.0/9.0, 1.0/9.0,
                                 1.0/36.0, 1.0/36.0, 1.0/36.0, 1.0/36.0};
    w3 = p->dim == 3 ? W3 : W2;
    for (int q = 0; q < c->Q; q++)
      for (size_t n = 0; n < nt; n++) h[q * nt + n] = w3[q] * p->rho0 * (1.0 + 0.0);
    *out = c;
    int rc = couplb_set_state_f(c, h.data());
    if (rc) return rc;
    std::vector<double> r(c->N, p->rho0);
    if (hipMemcpy(c->rho, r.data(), sizeof(double) * c->N, hipMemcpyHostToDevice) != hipSuccess)
      return fail(c, COUPLB_EDEVICE, "initial density upload failed");
    c->rho_fresh = 0;  // the first step's gravity reads the array (= rho0)
  }
  c->err.clear();
  return COUPLB_OK;
}

int couplb_destroy(couplb_ctx* c) {
  if (!c) return COUPLB_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  (void)hipFree(c->f[0]); (void)hipFree(c->f[1]); (void)hipFree(c->rho);
  (void)hipFree(c->mac); (void)hipFree(c->partial); (void)hipFree(c->red);
  (void)hipFree(c->u4); (void)hipFree(c->stamp);
  (void)hipFree(c->F4);
  (void)hipFree(c->flag); (void)hipFree(c->cover);
  (void)hipFree(c->keys); (void)hipFree(c->skeys); (void)hipFree(c->vals); (void)hipFree(c->perm);
  (void)hipFree(c->rowstart_cur); (void)hipFree(c->bases); (void)hipFree(c->sort_tmp);
  (void)hipFree(c->dx); (void)hipFree(c->dv); (void)hipFree(c->dm); (void)hipFree(c->df);
  (void)hipFree(c->ufc);
  (void)hipFree(c->staging);
  if (c->up_done) (void)hipEventDestroy(c->up_done);
  if (c->moved_ev) (void)hipEventDestroy(c->moved_ev);
  if (c->moved_host) (void)hipHostFree(c->moved_host);
  if (c->cstream) (void)hipStreamDestroy(c->cstream);
  for (int k = 0; k < HOST_CHUNKS_MAX; k++) {
    if (c->ev_up[k]) (void)hipEventDestroy(c->ev_up[k]);
    if (c->ev_cmp[k]) (void)hipEventDestroy(c->ev_cmp[k]);
  }
  if (c->dstream) (void)hipStreamDestroy(c->dstream);
  (void)hipFree(c->iperm);
  (void)hipFree(c->pkey); (void)hipFree(c->pkey_s); (void)hipFree(c->pval); (void)hipFree(c->pval_s);
  (void)hipFree(c->pp); (void)hipFree(c->pw); (void)hipFree(c->ptmp);
  (void)hipFree(c->pheads); (void)hipFree(c->phlen); (void)hipFree(c->phnode); (void)hipFree(c->phflag);
  (void)hipFree(c->pcount); (void)hipFree(c->seltmp); (void)hipFree(c->fsrc); (void)hipFree(c->plen);
  (void)hipFree(c->pseg);
  (void)hipFree(c->ut); (void)hipFree(c->mdf_v); (void)hipFree(c->mdf_S);
  (void)hipFree(c->mdf_F); (void)hipFree(c->mdf_R); (void)hipFree(c->mdf_Ap); (void)hipFree(c->mdf_part);
  (void)hipFree(c->mdf_Z);
  (void)hipFree(c->mlp_x); (void)hipFree(c->mlp_y); (void)hipFree(c->mlp_sig);
  (void)hipFree(c->mm_gam); (void)hipFree(c->mm_B);
  mm_free(c);
  (void)hipFree(c->mm_q);
  if (c->mdf_host) (void)hipHostFree(c->mdf_host);
  for (auto& e : c->mdf_ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& s : c->slots) (void)hipFree(s.rec);
  (void)hipFree(c->vcache); (void)hipFree(c->vbase);
  (void)hipFree(c->pcon); (void)hipFree(c->flag2); (void)hipFree(c->tflag);
  (void)hipFree(c->tlist);
  (void)hipFree(c->sb_cnt); (void)hipFree(c->sb_start); (void)hipFree(c->sb_bk); (void)hipFree(c->sb_rank);
  (void)hipFree(c->sb_ent); (void)hipFree(c->sb_ent2); (void)hipFree(c->sb_large); (void)hipFree(c->sb_scan_tmp);
  for (int e = 0; e < TimerRing::R; e++) {
    if (c->tr.a[e]) (void)hipEventDestroy(c->tr.a[e]);
    if (c->tr.b[e]) (void)hipEventDestroy(c->tr.b[e]);
  }
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return COUPLB_OK;
}

long couplb_state_len(const couplb_ctx* c) {
  if (!c) return 0;
  const long gx = c->geo.nx + 2, gy = c->geo.ny + 2, gz = c->D == 3 ? c->geo.nz + 2 : 1;
  return (long)c->Q * gx * gy * gz;
}

static int ensure_staging(couplb_ctx* c) {
  const size_t len = (size_t)couplb_state_len(c);
  if (c->staging_len < len) {
    (void)hipFree(c->staging);
    c->staging = nullptr;
    HIPCHK(c, hipMalloc(&c->staging, sizeof(double) * len));
    c->staging_len = len;
  }
  return COUPLB_OK;
}

int couplb_set_state_f(couplb_ctx* c, const double* fh) {
  if (!c || !fh) return fail(c, COUPLB_EINVAL, "null argument");
  int rc = ensure_staging(c);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->staging, fh, sizeof(double) * couplb_state_len(c),
                           hipMemcpyHostToDevice, c->stream));
  c->cur = 0;
  dispatch<LaunchImport>(c, (const double*)c->staging, c->f[0]);
  HIPCHK(c, hipGetLastError());
  c->local = 1;
  c->prev_valid = 0;
  c->nslots = 0;
  c->f4_partial = 0;
  c->cycle++;
  // read_checkpoint ends with compute_macroscopic(false) (couplb_io.h:817):
  // the density array then holds the restored density.
  rc = run_macro(c, true);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->rho, c->mac, sizeof(double) * c->N, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_init_equilibrium(couplb_ctx* c, double rho, double ux, double uy, double uz) {
  if (!c) return COUPLB_EINVAL;
  if (!(rho > 0) || !std::isfinite(ux) || !std::isfinite(uy) || !std::isfinite(uz))
    return fail(c, COUPLB_EINVAL, "init_equilibrium: rho must be positive and u finite");
  const double u[3] = {ux, uy, c->D == 3 ? uz : uz};
  c->cur = 0;
  dispatch<LaunchInitEq>(c, c->f[0], rho, u);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_fill, dim3((unsigned)((c->N + 255) / 256)), dim3(256), 0, c->stream, c->rho,
                     c->N, rho);
  HIPCHK(c, hipGetLastError());
  c->local = 1;
  c->prev_valid = 0;
  c->nslots = 0;
  c->f4_partial = 0;
  c->cycle++;
  c->rho_fresh = 0;  // like Grid::init_equilibrium: the array holds rho0
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_get_state_f(couplb_ctx* c, double* fh) {
  if (!c || !fh) return fail(c, COUPLB_EINVAL, "null argument");
  int rc = ensure_staging(c);
  if (rc) return rc;
  PullCtx pc = make_pull(c);
  // the other buffer holds the post-collision state of the step before the
  // last one (or, one step after the state was set, that state itself)
  const void* prev = c->prev_valid ? c->f[c->cur ^ 1] : nullptr;
  dispatch<LaunchExport>(c, pc, prev, c->staging);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(fh, c->staging, sizeof(double) * couplb_state_len(c),
                           hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_lbm_step(couplb_ctx* c, const double g[3], int flags) {
  if (!c) return COUPLB_EINVAL;
  return step_impl(c, g, flags, true);
}

int couplb_lbm_steps_async(couplb_ctx* c, const double g[3], int flags, int nsteps) {
  if (!c) return COUPLB_EINVAL;
  for (int s = 0; s < nsteps; s++) {
    // after the first step (which takes any spread IBM force, the import
    // form and a refreshed density), the rest of a small grid's run fuses
    if (s == 1 && nsteps > 2 && c->fused_lds > 0) return fused_steps(c, g, flags, nsteps - 1);
    const int rc = step_impl(c, g, flags, false);
    if (rc) return rc;
  }
  return COUPLB_OK;
}

void* couplb_host_alloc(size_t bytes) {
  void* p = nullptr;
  if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return p;
}

void couplb_host_free(void* p) {
  if (p) (void)hipHostFree(p);
}

int couplb_set_timer_mask(couplb_ctx* c, unsigned mask) {
  if (!c) return COUPLB_EINVAL;
  c->timer_mask = mask;
  return COUPLB_OK;
}

int couplb_lbm_fuses_steps(const couplb_ctx* c) { return (c && c->fused_lds > 0) ? 1 : 0; }

int couplb_ibm_substep_device(couplb_ctx* c, int n, const double* x, const double* v,
                              const double* m, double* f, const couplb_ibm_args* a) {
  if (!c) return COUPLB_EINVAL;
  c->hx_n = -1;  // the next host sub-step's "unchanged" refers to these positions
  return ibm_impl(c, n, x, v, m, f, 0, a);
}

int couplb_ibm_mdf_device(couplb_ctx* c, int n, const double* x, const double* V, const double* F0,
                          double* f_body, const couplb_mdf_args* a, const couplb_mdf_learn* learn,
                          couplb_mdf_stats* st) {
  if (!c) return COUPLB_EINVAL;
  c->hx_n = -1;
  return mdf_impl(c, n, x, V, F0, f_body, a, learn, st);
}

int couplb_ibm_substep(couplb_ctx* c, int n, const double* x, const double* v, const double* m,
                       double* f, const couplb_ibm_args* a) {
  if (!c || !a) return COUPLB_EINVAL;
  if (n <= 0) return COUPLB_OK;
  int rc = ensure_ibm(c, n);
  if (rc) return rc;
  if (!c->cstream) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->cstream, hipStreamNonBlocking));
    HIPCHK(c, hipEventCreateWithFlags(&c->up_done, hipEventDisableTiming));
  }
  const bool same = c->hx_n == n;  // dx, dm hold the previous host sub-step's x, m
  const IbmParams ip = make_ibm_params(c, a);
  if (same && n < (1 << REC_SLOT_SHIFT) && force_only(c, n, a->positions_unchanged, ip))
    return host_force_chunks(c, n, v, m, f, a, ip);
  // after an LB step the arrays go up on the copy stream, so that they
  // travel while its kernel runs (otherwise on the context stream, with no
  // cross-stream wait); positions and masses the caller declares unchanged
  // are already on the device from the previous host sub-step
  const bool overlap = c->cycle != c->host_cycle;
  hipStream_t up = overlap ? c->cstream : c->stream;
  if (!(a->positions_unchanged && same))
    HIPCHK(c, hipMemcpyAsync(c->dx, x, sizeof(double) * 3 * n, hipMemcpyHostToDevice, up));
  HIPCHK(c, hipMemcpyAsync(c->dv, v, sizeof(double) * 3 * n, hipMemcpyHostToDevice, up));
  if (!(a->masses_unchanged && same))
    HIPCHK(c, hipMemcpyAsync(c->dm, m, sizeof(double) * n, hipMemcpyHostToDevice, up));
  if (!a->forces_zero)
    HIPCHK(c, hipMemcpyAsync(c->df, f, sizeof(double) * 3 * n, hipMemcpyHostToDevice, up));
  if (overlap) {
    HIPCHK(c, hipEventRecord(c->up_done, c->cstream));
    HIPCHK(c, hipStreamWaitEvent(c->stream, c->up_done, 0));
  }
  c->hx_n = -1;
  c->host_cycle = c->cycle;
  // f += F on the device (fix_couplb.cpp:783-785), then back into f; with
  // f declared zero the device writes F itself
  rc = ibm_impl(c, n, c->dx, c->dv, c->dm, c->df, a->forces_zero ? 1 : 0, a);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(f, c->df, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->hx_n = n;
  return COUPLB_OK;
}

int couplb_refresh_macro(couplb_ctx* c) {
  if (!c) return COUPLB_EINVAL;
  c->rho_fresh = 1;
  return COUPLB_OK;
}

int couplb_get_macro(couplb_ctx* c, double* rho, double* ux, double* uy, double* uz) {
  if (!c) return COUPLB_EINVAL;
  int rc = run_macro(c, true);
  if (rc) return rc;
  double* dsts[4] = {rho, ux, uy, uz};
  for (int f = 0; f < 4; f++)
    if (dsts[f])
      HIPCHK(c, hipMemcpyAsync(dsts[f], c->mac + f * c->N, sizeof(double) * c->N,
                               hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_get_node_forces(couplb_ctx* c, double* f, int fold) {
  if (!c || !f) return COUPLB_EINVAL;
  const size_t N = c->N;
  if (c->nslots == 0 && !c->f4_partial) {
    std::memset(f, 0, sizeof(double) * 3 * N);
    return COUPLB_OK;
  }
  int rc = ensure_staging(c);
  if (rc) return rc;
  // staging (Q ghost-inclusive planes, at least 9 N doubles): 4 N force
  // records, N flag bytes, then the 3 N dense output
  double* F4x = c->staging;
  unsigned char* flagx = reinterpret_cast<unsigned char*>(c->staging + 4 * N);
  double* out = c->staging + 4 * N + (N + 7) / 8;
  if ((size_t)(out + 3 * N - c->staging) > c->staging_len)
    return fail(c, COUPLB_ESTATE, "staging buffer too small for the node forces");
  if (c->f4_partial) {
    // records of earlier sub-steps already spread (and folded) into F4
    HIPCHK(c, hipMemcpyAsync(F4x, c->F4, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(flagx, c->flag, N, hipMemcpyDeviceToDevice, c->stream));
  }
  if (c->nslots > 0) {
    rc = row_spread(c, F4x, flagx, fold ? 1 : 0, c->f4_partial);
    if (rc) return rc;
  }
  hipLaunchKernelGGL(k_force_export, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, c->stream, N,
                     (const double*)F4x, (const unsigned char*)flagx, out);
  HIPCHK(c, hipGetLastError());
  HIPCHK(c, hipMemcpyAsync(f, out, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_diagnostics(couplb_ctx* c, double out[6]) {
  if (!c || !out) return COUPLB_EINVAL;
  int rc = run_macro(c, false);
  if (rc) return rc;
  double h[6];
  HIPCHK(c, hipMemcpyAsync(h, c->red, sizeof(double) * 6, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = h[3];
  out[4] = std::sqrt(h[4]);
  out[5] = h[5];
  return COUPLB_OK;
}

int couplb_diagnostics_noalloc(couplb_ctx* c, double out[6]) {
  if (!c || !out) return fail(c, COUPLB_EINVAL, "null argument");
  if (!c->scratch) return fail(c, COUPLB_ESTATE, "no diagnostics scratch on this device");
  // every partial is written by its block: nothing to clear first
  MacroArgs a;
  a.pc = make_pull(c);
  a.rho = a.ux = a.uy = a.uz = nullptr;
  a.partial = c->scratch;
  dispatch<LaunchMacroParts>(c, a);
  HIPCHK(c, hipGetLastError());
  hipLaunchKernelGGL(k_macro_reduce, dim3(1), dim3(256), 0, c->stream, (const double*)c->scratch,
                     MACRO_PARTS, 1, c->scratch + SCRATCH_OUT, (double*)nullptr);
  HIPCHK(c, hipGetLastError());
  double h[6];
  HIPCHK(c, hipMemcpyAsync(h, c->scratch + SCRATCH_OUT, sizeof(double) * 6, hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->rho_fresh = 1;  // compute_macroscopic(false), as couplb_diagnostics
  out[0] = h[0]; out[1] = h[1]; out[2] = h[2]; out[3] = h[3];
  out[4] = std::sqrt(h[4]);
  out[5] = h[5];
  return COUPLB_OK;
}

int couplb_profile_y(couplb_ctx* c, double* rs, double* us, double* vs) {
  if (!c) return COUPLB_EINVAL;
  int rc = run_macro(c, false);
  if (rc) return rc;
  std::vector<double> h(3 * (size_t)c->geo.ny);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->red + 6, sizeof(double) * h.size(), hipMemcpyDeviceToHost,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int j = 0; j < c->geo.ny; j++) {
    if (rs) rs[j] = h[3 * j];
    if (us) us[j] = h[3 * j + 1];
    if (vs) vs[j] = h[3 * j + 2];
  }
  return COUPLB_OK;
}

int couplb_storage_saturations(couplb_ctx* c, unsigned long long* count) {
  if (!c || !count) return fail(c, COUPLB_EINVAL, "null argument");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  HIPCHK(c, hipMemcpyFromSymbol(count, HIP_SYMBOL(couplb::i16_saturations), sizeof(*count), 0,
                                hipMemcpyDeviceToHost));
  return COUPLB_OK;
}

int couplb_synchronize(couplb_ctx* c) {
  if (!c) return COUPLB_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return COUPLB_OK;
}

int couplb_copy_bandwidth(couplb_ctx* c, size_t bytes, int reps, double* gbs) {
  if (!c || !gbs || reps < 1 || bytes < 16) return COUPLB_EINVAL;
  const size_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = COUPLB_OK;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    c->err = "hipMalloc of the copy buffers failed";
    rc = COUPLB_ENOMEM;
  } else if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    c->err = "hipEventCreate failed";
    rc = COUPLB_EDEVICE;
  } else {
    const unsigned grid = (unsigned)((n + 256 * COPY_UNROLL - 1) / (256 * COPY_UNROLL));
    hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, (const CopyV*)a, (CopyV*)b, n);
    const hipError_t r0 = hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; r++)
      hipLaunchKernelGGL(k_copy16, dim3(grid), dim3(256), 0, c->stream, (const CopyV*)a, (CopyV*)b, n);
    const hipError_t r1 = hipEventRecord(e1, c->stream);
    float ms = 0.0f;
    if (r0 != hipSuccess || r1 != hipSuccess || hipGetLastError() != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.0f)) {
      c->err = "copy kernel failed";
      rc = COUPLB_EDEVICE;
    } else {
      *gbs = 2.0 * (double)(n * 16) * reps / (ms * 1e-3) / 1e9;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

void* couplb_stream(couplb_ctx* c) { return c ? (void*)c->stream : nullptr; }

double couplb_last_kernel_ms(couplb_ctx* c, int which) {
  if (!c || which < 0 || which >= TimerRing::W) return -1.0;
  if (timer_harvest_all(c)) return -1.0;
  return c->tr.last[which];
}
int couplb_kernel_ms_total(couplb_ctx* c, int which, double* total_ms, long* count) {
  if (!c || which < 0 || which >= TimerRing::W || !total_ms || !count) return COUPLB_EINVAL;
  const int rc = timer_harvest_all(c);
  if (rc) return rc;
  *total_ms = c->tr.total[which];
  *count = c->tr.count[which];
  return COUPLB_OK;
}
int couplb_kernel_ms_reset(couplb_ctx* c) {
  if (!c) return COUPLB_EINVAL;
  const int rc = timer_harvest_all(c);
  if (rc) return rc;
  for (int w = 0; w < TimerRing::W; w++) c->tr.total[w] = 0, c->tr.count[w] = 0;
  return COUPLB_OK;
}

}  // extern "C"
