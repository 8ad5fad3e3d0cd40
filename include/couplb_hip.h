/*
 * couplb_hip.h -- C ABI of the MI355X LB/IBM step (libcouplb_hip.so).
 *
 * This is the drop-in boundary under the LAMMPS fix: FixCoupLB's host code
 * (couplb_amd/host/fix_couplb_core.cpp) calls these functions instead of the
 * reference's header templates.  Each entry point names the reference code
 * it replaces.  Plain C types only; every function returns COUPLB_OK (0) or a
 * negative COUPLB_E* code and leaves a message for couplb_last_error().
 *
 * Units are lattice units (dx = dt_LB = 1), as inside the reference's Grid.
 * One context owns one grid on one GPU and one HIP stream; calls are
 * synchronous at return unless the name says _async.
 */
#ifndef COUPLB_HIP_H
#define COUPLB_HIP_H

#include <stddef.h> /* size_t */

#ifdef __cplusplus
extern "C" {
#endif

#define COUPLB_OK 0
#define COUPLB_EINVAL -1   /* bad argument or configuration */
#define COUPLB_EDEVICE -2  /* HIP runtime error */
#define COUPLB_ENOMEM -3   /* device allocation failed */
#define COUPLB_ESTATE -4   /* call not valid in the current state */

/* population storage (precision variants) */
#define COUPLB_FP64 0 /* f_i, the reference's representation */
#define COUPLB_FP32 1 /* f_i - w_i*rho0 in float, float arithmetic */
#define COUPLB_FP16 2 /* f_i - w_i*rho0 in IEEE half, float arithmetic */
#define COUPLB_BF16 3 /* f_i - w_i*rho0 in bfloat16, float arithmetic */
#define COUPLB_I16 4  /* f_i - w_i*rho0 in 16-bit fixed point (storage_range),
                         float arithmetic, node mass kept exactly */

/* IBM delta kernels, couplb_ibm.h:11-33 / fix_couplb.cpp:175-180 */
#define COUPLB_ROMA 0
#define COUPLB_PESKIN4 1

/* face types, fix_couplb.cpp:49-51 */
#define COUPLB_PERIODIC 0
#define COUPLB_NOSLIP 1
#define COUPLB_MOVING 2
#define COUPLB_FREESLIP 3
#define COUPLB_OPEN 4

typedef struct couplb_ctx couplb_ctx;

/* Grid configuration.  Replaces Grid<L>::allocate + init_equilibrium
 * (couplb_lattice.h:105-154), BGK<L>::set_tau (couplb_collision.h:16),
 * Boundary<L>::set_walls_x/y/z + set_wall_velocity (couplb_boundary.h:12-38)
 * and Streaming<L>::allocate_buffers/precompute_wall_flags
 * (couplb_streaming.h:54-80). */
typedef struct couplb_params {
  int dim;                /* 2 -> D2Q9, 3 -> D3Q19 */
  int nx, ny, nz;         /* interior nodes; nz = 1 when dim = 2 */
  double tau;             /* BGK relaxation time, > 0.5 */
  int wall[3][2];         /* COUPLB_PERIODIC..COUPLB_OPEN for (lo, hi) of x, y, z */
  double face_vel[3][2][3]; /* velocity of each type-2 face; the reference
                               gives every face the same one */
  int kernel;             /* COUPLB_ROMA or COUPLB_PESKIN4 */
  int storage;            /* COUPLB_FP64 .. COUPLB_I16 */
  double rho0;            /* lattice reference density (fix_couplb.h:41: 1.0) */
  int md_per_lb;          /* IBM sub-steps per LB step (sizes record slots) */
  int device;             /* HIP device ordinal */
  double storage_range;   /* COUPLB_I16: largest |f_i - w_i*rho0| to hold
                             (the resolution is the largest power of two
                             with 32767 steps inside it); 0: 1/32 */
} couplb_params;

int couplb_create(const couplb_params* p, couplb_ctx** out);
int couplb_destroy(couplb_ctx* ctx);
const char* couplb_last_error(const couplb_ctx* ctx);
const char* couplb_version(void);

/* Population state in the reference's ghost-inclusive layout
 * f[q*ntotal + n], n = i + gx*(j + gy*k), gx = nx+2, gy = ny+2,
 * gz = nz+2 (3D) or 1 (2D) -- exactly the array couplb_io.h:653-828 writes.
 * set: replaces read_checkpoint's `f = file; f_buf = f;
 * compute_macroscopic(false)` (couplb_io.h:802-817).  Ghost entries are
 * ignored on input.  get: returns the whole f the reference holds between
 * steps, which its checkpoint writes (couplb_io.h:706): post-stream
 * populations inside; on periodic ghost planes the images of the
 * post-collision populations of the step before the last one (what
 * Streaming::exchange left there before the swap, couplb_streaming.h:
 * 133-182), or of the current state when no step has run since it was set;
 * on wall ghost planes the init_equilibrium value w_q*rho0.  One step after
 * a couplb_set_state_f the reference's ghosts are the imported file's, which
 * the device does not keep: it returns the images of the imported state. */
int couplb_set_state_f(couplb_ctx* ctx, const double* f_host);
int couplb_get_state_f(couplb_ctx* ctx, double* f_host);
long couplb_state_len(const couplb_ctx* ctx); /* Q*ntotal */

/* Uniform equilibrium state f_q = f^eq_q(rho, u) at every node and the
 * density array set to rho: replaces Grid::init_equilibrium
 * (couplb_lattice.h:145-154), which the fix calls with (1, 0, 0, 0) at
 * fix_couplb.cpp:504/513.  couplb_create already applies (rho0, 0, 0, 0). */
int couplb_init_equilibrium(couplb_ctx* ctx, double rho, double ux, double uy, double uz);

/* One LB step: replaces do_lbm_step (fix_couplb.cpp:659-669), i.e.
 * apply_external_force -> compute_macroscopic(true) -> BGK::collide ->
 * Streaming::exchange -> Streaming::stream -> clear_forces, fused into one
 * pull kernel.  g_lb is the lattice-unit body acceleration
 * (fix_couplb.cpp:688-691); pass {0,0,0} for none.  IBM forces are the
 * records of every couplb_ibm_substep since the previous step.
 * flags: COUPLB_STEP_RHO_REFRESHED_NEXT -- the caller guarantees a
 * compute_macroscopic(false)-equivalent call (an IBM sub-step with markers,
 * couplb_refresh_macro, couplb_diagnostics) before the next step, so the
 * density array need not be kept for the next step's gravity. */
#define COUPLB_STEP_RHO_REFRESHED_NEXT 1
/* Asynchronous on the context stream; calls that return data to the host
 * (get_state_f, get_macro, diagnostics, profile_y, the host-array IBM
 * sub-step) wait for it. */
int couplb_lbm_step(couplb_ctx* ctx, const double g_lb[3], int flags);
/* nsteps steps with the same gravity and flags.  On a grid of at most
 * 1024 nodes (COUPLB_FUSED_NODES overrides the bound) whose two states fit
 * the LDS of one workgroup, every step after the first runs inside one
 * launch of one workgroup, with the same results. */
int couplb_lbm_steps_async(couplb_ctx* ctx, const double g_lb[3], int flags, int nsteps);
/* 1 when couplb_lbm_steps_async runs the steps of this grid in one launch.
 * No counterpart in the reference (its steps are host loops). */
int couplb_lbm_fuses_steps(const couplb_ctx* ctx);
/* Timed kinds (bit k for kind k of couplb_kernel_ms_total; all by default).
 * Each timed region records two HIP events on the stream, which costs a few
 * microseconds between kernels; untimed kinds record none.  Measurement
 * helper, no counterpart in the reference. */
int couplb_set_timer_mask(couplb_ctx* ctx, unsigned mask);
/* Pinned (page-locked) host memory for the caller's packing buffers, so
 * the host-array IBM sub-step's copies run as DMA; null on failure.  No
 * counterpart in the reference (its arrays never leave the host). */
void* couplb_host_alloc(size_t bytes);
void couplb_host_free(void* p);

/* IBM sub-step: replaces the loop body of do_ibm_sub_coupling
 * (fix_couplb.cpp:757-797): fluid velocity at the marker (IBM::interpolate,
 * couplb_ibm.h:57-90), penalty force
 *   F = m * xi_over_dtLB * (u_f * vel_scale - v)        (779-781)
 * added into f, and the reaction -F*inv_N/force_scale spread onto the grid
 * (IBM::spread, couplb_ibm.h:92-133) with the self-periodic fold
 * (couplb_streaming.h:217-225) applied as an index wrap.  The fluid velocity
 * is compute_macroscopic(false) of the current state, as at 757-760; it is
 * computed once per state and reused across sub-steps.
 * Arrays are [n][3] (x, v, f) and [n] (m) -- LAMMPS' own layout. */
typedef struct couplb_ibm_args {
  double lo[3];          /* domain->boxlo (fix_couplb.cpp:282) */
  double inv_dx;         /* 1/dx */
  double xi_over_dtLB;   /* xi_ibm / dt_LBM (fix_couplb.cpp:763) */
  double vel_scale;      /* dx / dt_LBM */
  double inv_N;          /* 1 / md_per_lb */
  double force_scale;    /* rho*dx^(D+1)/dt_LBM^2 */
  int positions_unchanged; /* 1: x equals the previous sub-step's x, so the
                              marker ordering may be reused */
  int masses_unchanged;    /* couplb_ibm_substep only (0 elsewhere): 1 when m
                              equals the previous host sub-step's m, which is
                              then not sent again */
  int forces_zero;         /* couplb_ibm_substep only (0 elsewhere): 1 when
                              f_inout holds zeros on entry, so it returns F
                              without sending f to the device */
} couplb_ibm_args;

/* host arrays (copied over PCIe each call; x and m only when the flags above
 * do not declare them unchanged, f only when not declared zero; the uploads
 * run on a copy stream of their own, so after couplb_lbm_step they travel
 * while the LB kernel runs) */
int couplb_ibm_substep(couplb_ctx* ctx, int n, const double* x, const double* v,
                       const double* m, double* f_inout, const couplb_ibm_args* a);
/* device arrays (already resident in HBM); f_inout is device memory.
 * Asynchronous on the context stream: the arrays must stay valid until the
 * stream has passed this call (couplb_synchronize). */
int couplb_ibm_substep_device(couplb_ctx* ctx, int n, const double* x, const double* v,
                              const double* m, double* f_inout, const couplb_ibm_args* a);

/* Multi-direct-forcing IBM for markers with prescribed velocity.  No
 * counterpart in the reference, which couples with one penalty pass per
 * sub-step (fix_couplb.cpp:779-785).  Iterates the marker forces F (force on
 * the fluid, lattice units) until the velocity of the next collision,
 * (j + F/2)/rho, interpolates to V at every marker:
 *   slip(F) = V/vel_scale - I(u* + S(F)/(2 rho)) = 0   to max|slip| < tol,
 * with u* = j/rho of the current state, I = IBM::interpolate and
 * S = IBM::spread (couplb_ibm.h:57-133), by Richardson steps
 * F += 2*relax*slip (classic multi-direct forcing) or by conjugate gradients,
 * plain or block-Jacobi preconditioned (restarted from the true slip; falls
 * back to Richardson if it breaks down).  The converged F is spread into the
 * next LB step.  It must be the only IBM call between two LB steps.
 * x, V (physical units), F0 (initial F in lattice units, caller order; NULL
 * starts from zero) and f_body are device arrays [n][3]; f_body receives the
 * force on the body per marker, -F*force_scale.  Iterations run on the
 * device and stop on a device-side flag; the host checks it every few
 * iterations.  stats->iterations counts interpolation passes. */
#define COUPLB_MDF_RICHARDSON 0  /* F += 2*relax*slip, the classic iteration */
#define COUPLB_MDF_CG 1          /* conjugate gradients on the same fixed point */
#define COUPLB_MDF_PCG 2         /* conjugate gradients preconditioned by the inverses
                                    of 32-marker blocks of neighbouring markers (on the
                                    explicit operator; plain CG where it is not built) */
typedef struct couplb_mdf_args {
  couplb_ibm_args ibm;  /* lo, inv_dx, vel_scale, force_scale, positions_unchanged */
  double tol;           /* slip tolerance, lattice velocity units */
  double relax;         /* omega of the Richardson iteration */
  int max_iter;         /* operator applications (interpolation passes) */
  int method;           /* COUPLB_MDF_RICHARDSON, COUPLB_MDF_CG or COUPLB_MDF_PCG */
} couplb_mdf_args;
typedef struct couplb_mdf_stats {
  int iterations;       /* interpolation passes, including the one at F = 0
                           when the network is used */
  int converged;        /* 1 when max|slip| < tol was reached */
  double slip_first;    /* max|slip| at the force the iteration starts from */
  double slip_last;     /* max|slip| of the returned force */
  double slip_zero;     /* max|slip| at F = 0 when evaluated, else -1 */
} couplb_mdf_stats;

/* Learned warm start (include/couplb_mlp.h).  The network's inputs for
 * marker k come from the slip at F = 0, s = V/vel_scale - I(u*), and the
 * velocities of its 4x4x4 stencil: 64 values ((V - u_l).s)/|s|^2 in stencil
 * order t + 4u + 16v (u_l = u* at fluid nodes, the wall velocity at moving
 * walls, 0 at other walls and outside the ghost layer) and s/|s|; its 3
 * outputs are F/(16 |s|).  warm_start: start the iteration from the
 * network's prediction (instead of F0); features / targets (device, caller
 * order, optional): the inputs of this solve and the converged forces
 * scaled as outputs, for training. */
struct couplb_mlp;
typedef struct couplb_mdf_learn {
  struct couplb_mlp* warm_start;  /* may be NULL */
  float* features;                /* [n][67] or NULL */
  float* targets;                 /* [n][3] or NULL */
} couplb_mdf_learn;

int couplb_ibm_mdf_device(couplb_ctx* ctx, int n, const double* x, const double* V,
                          const double* F0, double* f_body, const couplb_mdf_args* a,
                          const couplb_mdf_learn* learn, couplb_mdf_stats* stats);

/* compute_macroscopic(false) semantics without copying anything out: marks
 * the density as refreshed for the next step's gravity (fix_couplb.cpp:706,
 * couplb_lattice.h:155-202). */
int couplb_refresh_macro(couplb_ctx* ctx);

/* Interior macroscopic fields (rho, u = j/rho) of the current state, each an
 * nx*ny*nz array with x fastest.  Also a refresh. */
int couplb_get_macro(couplb_ctx* ctx, double* rho, double* ux, double* uy, double* uz);

/* Node forces of the IBM reactions recorded since the last LB step, as the
 * reference's Grid::fx/fy/fz hold them at end_of_step (lattice units): the
 * records spread by IBM::spread (couplb_ibm.h:92-133); fold = 1 adds the
 * periodic ghost images into the interior as exchange_forces does after the
 * interval's last sub-step (couplb_streaming.h:217-225), fold = 0 keeps the
 * direct contributions only.  f: 3*nx*ny*nz doubles, (fx, fy, fz) per node,
 * x fastest.  Feeds the VTK force field (couplb_io.h:255-258). */
int couplb_get_node_forces(couplb_ctx* ctx, double* f, int fold);

/* FP64 global diagnostics: replaces Grid::max_velocity and
 * Grid::compute_diagnostics (couplb_lattice.h:203-242) used by
 * check_stability_precomputed (fix_couplb.cpp:818-844).
 * out = {mass, momx, momy, momz, umax, clamp_count}.  Also a refresh. */
int couplb_diagnostics(couplb_ctx* ctx, double out[6]);

/* couplb_diagnostics without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same out[6], bit for bit.  Its partial sums live in a fixed 18,496-byte
 * array of device memory that the library holds on every GPU it runs on
 * (256 block partials whatever the grid size, so no grid is too large); the
 * array is shared by every context on that GPU and every call rewrites all
 * of it, so calls that overlap on one GPU (contexts on different streams or
 * host threads) race.  Also a refresh. */
int couplb_diagnostics_noalloc(couplb_ctx* ctx, double out[6]);

/* y-binned sums over x and z of rho, ux, uy (write_profile,
 * fix_couplb.cpp:903-928); each array has ny entries.  Also a refresh. */
int couplb_profile_y(couplb_ctx* ctx, double* rho_sum, double* ux_sum, double* uy_sum);

/* Populations the 16-bit fixed-point storage (COUPLB_I16) had to clamp to
 * its range on this context's GPU since the library was loaded (all
 * contexts of that GPU together).  No counterpart in the reference. */
int couplb_storage_saturations(couplb_ctx* ctx, unsigned long long* count);

int couplb_synchronize(couplb_ctx* ctx);
/* Measurement helper, no counterpart in the reference: read + write GB/s of
 * a plain 16-byte-per-lane copy of `bytes` between two fresh device buffers
 * (allocated and freed here), timed over `reps` launches on the context
 * stream after one untimed launch. */
int couplb_copy_bandwidth(couplb_ctx* ctx, size_t bytes, int reps, double* gbs);
/* the context's hipStream_t, for callers that time with HIP events */
void* couplb_stream(couplb_ctx* ctx);

/* Kernel timing from HIP events on the context stream, read back without
 * stalling the stream: which = 0 fused LB kernel, 1 IBM sub-step (all its
 * kernels), 2 force spreading before an LB kernel, 3 MDF solve, 4 the fluid
 * velocity gather at covered nodes (part of 1 or 3; once per LB state and
 * marker ordering).
 * couplb_last_kernel_ms waits for the newest region of that kind and returns
 * its milliseconds; couplb_kernel_ms_total returns the sum and count of all
 * regions since the last reset. */
double couplb_last_kernel_ms(couplb_ctx* ctx, int which);
int couplb_kernel_ms_total(couplb_ctx* ctx, int which, double* total_ms, long* count);
int couplb_kernel_ms_reset(couplb_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
