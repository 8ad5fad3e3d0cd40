/*
 * couplb_oracle.c -- CPU FP64 restatement of the CoupLB LB/IBM step.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle for the HIP path; see the header).
 * Built with -ffp-contract=off so every expression rounds as the reference
 * does when compiled the same way; expression order follows the reference so
 * that results agree bit-for-bit with oracle/_ref (the reference headers).
 */
#include "couplb_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants, couplb_lattice.h:13-20 */
#define RHO_MIN        1e-10
#define DELTA_TOL      1e-12
#define ZERO_TOL       1e-15
#define RHO_CLAMP_FRAC 0.1
#define MA_WARN        0.3
#define MA_ERROR       0.5

/* ------------------------------------------------------------------ */
/* Lattice tables.  The velocity order, weights, opposite and reflection
 * tables must equal couplb_lattice.h:24-75 because populations are stored
 * and exchanged by index.                                              */
/* ------------------------------------------------------------------ */
typedef struct {
  int D, Q;
  int e[19][3];
  double w[19];
  int opp[19];
  int reflect[3][19];
} lattice_t;

static const lattice_t L2 = {
  2, 9,
  {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0}},
  {4.0/9.0, 1.0/9.0,1.0/9.0,1.0/9.0,1.0/9.0, 1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0},
  {0,2,1,4,3,6,5,8,7},
  {{0,2,1,3,4,8,7,6,5},{0,1,2,4,3,7,8,5,6},{0,1,2,3,4,5,6,7,8}}
};

static const lattice_t L3 = {
  3, 19,
  {{0,0,0},{1,0,0},{-1,0,0},{0,1,0},{0,-1,0},{0,0,1},{0,0,-1},
   {1,1,0},{-1,-1,0},{1,-1,0},{-1,1,0},
   {1,0,1},{-1,0,-1},{1,0,-1},{-1,0,1},
   {0,1,1},{0,-1,-1},{0,1,-1},{0,-1,1}},
  {1.0/3.0, 1.0/18.0,1.0/18.0,1.0/18.0,1.0/18.0,1.0/18.0,1.0/18.0,
   1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0,
   1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0,1.0/36.0},
  {0,2,1,4,3,6,5,8,7,10,9,12,11,14,13,16,15,18,17},
  {{0,2,1,3,4,5,6,10,9,8,7,14,13,12,11,15,16,17,18},
   {0,1,2,4,3,5,6,9,10,7,8,11,12,13,14,18,17,16,15},
   {0,1,2,3,4,6,5,7,8,9,10,13,14,11,12,17,18,15,16}}
};

static const double CS2 = 1.0 / 3.0;   /* Lattice::cs2 */

/* ------------------------------------------------------------------ */
/* Grid (couplb_lattice.h:82-248)                                      */
/* ------------------------------------------------------------------ */
typedef struct {
  const lattice_t* L;
  int D, Q, nx, ny, nz, gx, gy, gz;
  long ntotal;
  double *f, *f_buf, *rho, *ux, *uy, *uz, *fx, *fy, *fz, *bcx, *bcy, *bcz;
  int *type, *wall_dim;
  int clamp_count;
  double omega;
  int wall_face[3][2];
  double rho0;
} grid_t;

static inline long gidx(const grid_t* g, int i, int j, int k) {
  return (long)i + (long)g->gx * ((long)j + (long)g->gy * (long)k);
}

static double* dalloc(long n, double v) {
  double* p = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (long i = 0; i < n; i++) p[i] = v;
  return p;
}

static inline double feq_of(const lattice_t* L, int q, double r,
                            double u0, double u1, double u2) {
  /* Grid::feq, couplb_lattice.h:137-144 */
  const double eu = L->e[q][0]*u0 + L->e[q][1]*u1 + L->e[q][2]*u2;
  const double uu = u0*u0 + u1*u1 + u2*u2;
  const double cs4 = CS2 * CS2;
  return L->w[q] * r * (1.0 + eu/CS2 + 0.5*eu*eu/cs4 - 0.5*uu/CS2);
}

/* Boundary::set_walls_x/y/z (couplb_boundary.h:12-35) followed by the
 * per-face velocity assignment.  The reference assigns one velocity to every
 * type-2 node (couplb_boundary.h:36-38, fix_couplb.cpp:532); with all faces
 * given the same velocity this is identical. */
static void set_walls(grid_t* g, const clo_grid_params* p) {
  for (int d = 0; d < g->D; d++) {
    for (int s = 0; s < 2; s++) {
      const int t = p->wall[d][s];
      if (t <= 0) continue;
      const int pos = (s == 0) ? 0 : ((d == 0) ? g->gx - 1 : (d == 1) ? g->gy - 1 : g->gz - 1);
      for (int k = 0; k < g->gz; k++)
        for (int j = 0; j < g->gy; j++)
          for (int i = 0; i < g->gx; i++) {
            const int c[3] = {i, j, k};
            if (c[d] != pos) continue;
            const long n = gidx(g, i, j, k);
            g->type[n] = t;
            g->wall_dim[n] = d;
            if (t == 2) {
              g->bcx[n] = p->face_vel[d][s][0];
              g->bcy[n] = p->face_vel[d][s][1];
              g->bcz[n] = (g->D == 3) ? p->face_vel[d][s][2] : 0.0;
            }
          }
    }
  }
}

/* Streaming::has_wall_on_face / precompute_wall_flags, couplb_streaming.h:74-80, 163-169 */
static int face_is_wall(const grid_t* g, int d, int pos) {
  for (int k = 0; k < g->gz; k++)
    for (int j = 0; j < g->gy; j++)
      for (int i = 0; i < g->gx; i++) {
        const int c[3] = {i, j, k};
        if (c[d] != pos) continue;
        if (g->type[gidx(g, i, j, k)] == 0) return 0;
      }
  return 1;
}

static void* o_create(const clo_grid_params* p) {
  grid_t* g = (grid_t*)calloc(1, sizeof(grid_t));
  g->L = (p->D == 3) ? &L3 : &L2;
  g->D = p->D; g->Q = g->L->Q;
  g->nx = p->nx; g->ny = p->ny; g->nz = (p->D == 3) ? p->nz : 1;
  g->gx = g->nx + 2; g->gy = g->ny + 2; g->gz = (p->D == 3) ? g->nz + 2 : 1;
  g->ntotal = (long)g->gx * g->gy * g->gz;
  const long nt = g->ntotal;
  g->f = dalloc(g->Q * nt, 0.0);
  g->f_buf = dalloc(g->Q * nt, 0.0);
  g->rho = dalloc(nt, 1.0);
  g->ux = dalloc(nt, 0.0); g->uy = dalloc(nt, 0.0); g->uz = dalloc(nt, 0.0);
  g->fx = dalloc(nt, 0.0); g->fy = dalloc(nt, 0.0); g->fz = dalloc(nt, 0.0);
  g->bcx = dalloc(nt, 0.0); g->bcy = dalloc(nt, 0.0); g->bcz = dalloc(nt, 0.0);
  g->type = (int*)calloc((size_t)nt, sizeof(int));
  g->wall_dim = (int*)malloc(sizeof(int) * (size_t)nt);
  for (long n = 0; n < nt; n++) g->wall_dim[n] = -1;
  g->omega = 1.0 / p->tau;
  g->rho0 = p->rho_lb;
  /* init_equilibrium(rho_lb, 0, 0, 0), couplb_lattice.h:145-154 */
  for (long n = 0; n < nt; n++) {
    g->rho[n] = p->rho_lb;
    for (int q = 0; q < g->Q; q++)
      g->f[q * nt + n] = feq_of(g->L, q, p->rho_lb, 0.0, 0.0, 0.0);
  }
  memcpy(g->f_buf, g->f, sizeof(double) * (size_t)(g->Q * nt));
  set_walls(g, p);
  for (int d = 0; d < g->D; d++) {
    const int ext = (d == 0) ? g->gx : (d == 1) ? g->gy : g->gz;
    g->wall_face[d][0] = face_is_wall(g, d, 0);
    g->wall_face[d][1] = face_is_wall(g, d, ext - 1);
  }
  return g;
}

static void o_destroy(void* vg) {
  grid_t* g = (grid_t*)vg;
  if (!g) return;
  free(g->f); free(g->f_buf); free(g->rho); free(g->ux); free(g->uy); free(g->uz);
  free(g->fx); free(g->fy); free(g->fz); free(g->bcx); free(g->bcy); free(g->bcz);
  free(g->type); free(g->wall_dim); free(g);
}

static double* o_dfield(void* vg, int which) {
  grid_t* g = (grid_t*)vg;
  switch (which) {
    case CLO_F_RHO: return g->rho;  case CLO_F_UX: return g->ux;
    case CLO_F_UY: return g->uy;    case CLO_F_UZ: return g->uz;
    case CLO_F_FX: return g->fx;    case CLO_F_FY: return g->fy;
    case CLO_F_FZ: return g->fz;    case CLO_F_BCUX: return g->bcx;
    case CLO_F_BCUY: return g->bcy; case CLO_F_BCUZ: return g->bcz;
    case CLO_F_F: return g->f;      case CLO_F_FBUF: return g->f_buf;
  }
  return NULL;
}
static int* o_ifield(void* vg, int which) {
  grid_t* g = (grid_t*)vg;
  return which == CLO_I_TYPE ? g->type : g->wall_dim;
}
static void o_dims(void* vg, int gd[3], long* ntotal, int* Q) {
  grid_t* g = (grid_t*)vg;
  gd[0] = g->gx; gd[1] = g->gy; gd[2] = g->gz; *ntotal = g->ntotal; *Q = g->Q;
}

/* Grid::compute_macroscopic, couplb_lattice.h:155-202 */
static int o_macro(void* vg, int guo) {
  grid_t* g = (grid_t*)vg;
  const lattice_t* L = g->L;
  const long nt = g->ntotal;
  const int klo = (g->D == 3) ? 1 : 0, khi = (g->D == 3) ? g->gz - 2 : 0;
  int clamps = 0;
#pragma omp parallel for collapse(2) schedule(static) reduction(+:clamps)
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= g->gy - 2; j++)
      for (int i = 1; i <= g->gx - 2; i++) {
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        double r = 0.0, su = 0.0, sv = 0.0, sw = 0.0;
        for (int q = 0; q < L->Q; q++) {
          const double fv = g->f[q * nt + n];
          r += fv;
          su += L->e[q][0] * fv;
          sv += L->e[q][1] * fv;
          sw += L->e[q][2] * fv;
        }
        if (r <= RHO_MIN) { r = RHO_MIN; clamps++; }
        if (!isfinite(r)) { r = RHO_MIN; clamps++; }
        g->rho[n] = r;
        const double inv_r = 1.0 / r;
        if (guo) {
          g->ux[n] = (su + 0.5 * g->fx[n]) * inv_r;
          g->uy[n] = (sv + 0.5 * g->fy[n]) * inv_r;
          g->uz[n] = (sw + 0.5 * g->fz[n]) * inv_r;
        } else {
          g->ux[n] = su * inv_r;
          g->uy[n] = sv * inv_r;
          g->uz[n] = sw * inv_r;
        }
      }
  g->clamp_count = clamps;
  return clamps;
}

/* BGK::collide, couplb_collision.h:18-62 */
static void o_collide(void* vg) {
  grid_t* g = (grid_t*)vg;
  const lattice_t* L = g->L;
  const long nt = g->ntotal;
  const double cs4 = CS2 * CS2;
  const double om = g->omega;
  const double half_om = 1.0 - 0.5 * om;
  const int klo = (g->D == 3) ? 1 : 0, khi = (g->D == 3) ? g->gz - 2 : 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= g->gy - 2; j++)
      for (int i = 1; i <= g->gx - 2; i++) {
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        const double r = g->rho[n];
        const double u0 = g->ux[n], u1 = g->uy[n], u2 = g->uz[n];
        const double f0 = g->fx[n], f1 = g->fy[n], f2 = g->fz[n];
        const double uu = u0*u0 + u1*u1 + u2*u2;
        const double uf = u0*f0 + u1*f1 + u2*f2;
        for (int q = 0; q < L->Q; q++) {
          const double cx = L->e[q][0], cy = L->e[q][1], cz = L->e[q][2];
          const double w = L->w[q];
          const double eu = cx*u0 + cy*u1 + cz*u2;
          const double ef = cx*f0 + cy*f1 + cz*f2;
          const double fe = w * r * (1.0 + eu/CS2 + 0.5*eu*eu/cs4 - 0.5*uu/CS2);
          const double Si = w * half_om * ((ef - uf)/CS2 + eu*ef/cs4);
          double* fp = &g->f[q * nt + n];
          *fp += -om * (*fp - fe) + Si;
        }
      }
}

/* Face iteration over the full ghost-inclusive extent of the other two
 * dimensions (couplb_streaming.h:281-293). */
#define FACE_LOOP(g, d, pos, n, BODY)                                    \
  do {                                                                   \
    if ((d) == 0) { for (int k_ = 0; k_ < (g)->gz; k_++) for (int j_ = 0; j_ < (g)->gy; j_++) { const long n = gidx(g, pos, j_, k_); BODY; } } \
    else if ((d) == 1) { for (int k_ = 0; k_ < (g)->gz; k_++) for (int i_ = 0; i_ < (g)->gx; i_++) { const long n = gidx(g, i_, pos, k_); BODY; } } \
    else { for (int j_ = 0; j_ < (g)->gy; j_++) for (int i_ = 0; i_ < (g)->gx; i_++) { const long n = gidx(g, i_, j_, pos); BODY; } } \
  } while (0)

static int dim_ext(const grid_t* g, int d) { return d == 0 ? g->gx : d == 1 ? g->gy : g->gz; }

/* Paired face copy: for every node on plane `s` copy to the node on plane
 * `t` with the same other two coordinates. */
static void copy_face_f(grid_t* g, int d, int s, int t) {
  const long nt = g->ntotal;
  const long stride = (d == 0) ? 1 : (d == 1) ? g->gx : (long)g->gx * g->gy;
  FACE_LOOP(g, d, s, n, {
    const long dn = n + (long)(t - s) * stride;
    for (int q = 0; q < g->Q; q++) g->f[q * nt + dn] = g->f[q * nt + n];
  });
}
static void copy_face_vel(grid_t* g, int d, int s, int t) {
  const long stride = (d == 0) ? 1 : (d == 1) ? g->gx : (long)g->gx * g->gy;
  FACE_LOOP(g, d, s, n, {
    const long dn = n + (long)(t - s) * stride;
    g->rho[dn] = g->rho[n]; g->ux[dn] = g->ux[n]; g->uy[dn] = g->uy[n]; g->uz[dn] = g->uz[n];
  });
}
static void accum_face_force(grid_t* g, int d, int s, int t) {
  const long stride = (d == 0) ? 1 : (d == 1) ? g->gx : (long)g->gx * g->gy;
  FACE_LOOP(g, d, s, n, {
    const long dn = n + (long)(t - s) * stride;
    g->fx[dn] += g->fx[n]; g->fy[dn] += g->fy[n]; g->fz[dn] += g->fz[n];
  });
}

/* Streaming::exchange (self-periodic branch), couplb_streaming.h:136-139, 176-182.
 * A face without a wall takes the periodic image; a single-rank grid has
 * every dimension self-neighboured, so this is the only branch exercised. */
static void o_exchange(void* vg) {
  grid_t* g = (grid_t*)vg;
  for (int d = 0; d < g->D; d++) {
    const int ds = dim_ext(g, d);
    if (ds < 4) continue;
    if (!g->wall_face[d][0]) copy_face_f(g, d, ds - 2, 0);
    if (!g->wall_face[d][1]) copy_face_f(g, d, 1, ds - 1);
  }
}
/* Streaming::exchange_velocity, couplb_streaming.h:140-143, 196-202 */
static void o_exchange_velocity(void* vg) {
  grid_t* g = (grid_t*)vg;
  for (int d = 0; d < g->D; d++) {
    const int ds = dim_ext(g, d);
    if (ds < 4) continue;
    if (!g->wall_face[d][0]) copy_face_vel(g, d, ds - 2, 0);
    if (!g->wall_face[d][1]) copy_face_vel(g, d, 1, ds - 1);
  }
}
/* Streaming::exchange_forces (self-periodic), couplb_streaming.h:155-158, 217-225:
 * ghost lo -> interior hi face, ghost hi -> interior lo face. */
static void o_exchange_forces(void* vg) {
  grid_t* g = (grid_t*)vg;
  for (int d = 0; d < g->D; d++) {
    const int ds = dim_ext(g, d);
    if (ds < 4) continue;
    accum_face_force(g, d, 0, ds - 2);
    accum_face_force(g, d, ds - 1, 1);
  }
}

/* Streaming::stream, couplb_streaming.h:82-134 */
static void o_stream(void* vg) {
  grid_t* g = (grid_t*)vg;
  const lattice_t* L = g->L;
  const long nt = g->ntotal;
  const double rho_floor = (RHO_CLAMP_FRAC * g->rho0 > RHO_MIN) ? RHO_CLAMP_FRAC * g->rho0 : RHO_MIN;
  const int klo = (g->D == 3) ? 1 : 0, khi = (g->D == 3) ? g->gz - 2 : 0;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= g->gy - 2; j++)
      for (int i = 1; i <= g->gx - 2; i++) {
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        for (int q = 0; q < L->Q; q++) {
          const int si = i - L->e[q][0], sj = j - L->e[q][1], sk = k - L->e[q][2];
          double* out = &g->f_buf[q * nt + n];
          if (si < 0 || si >= g->gx || sj < 0 || sj >= g->gy ||
              (g->D == 3 && (sk < 0 || sk >= g->gz))) {
            *out = g->f[q * nt + n];
            continue;
          }
          const long sn = gidx(g, si, sj, sk);
          const int st = g->type[sn];
          if (st == 0) {
            *out = g->f[q * nt + sn];
          } else if (st == 1) {
            *out = g->f[L->opp[q] * nt + n];
          } else if (st == 2) {
            const int qo = L->opp[q];
            const double eu = L->e[q][0]*g->bcx[sn] + L->e[q][1]*g->bcy[sn] + L->e[q][2]*g->bcz[sn];
            const double rw = (g->rho[n] < rho_floor) ? rho_floor : g->rho[n];
            *out = g->f[qo * nt + n] + 2.0*L->w[qo]*rw*eu/CS2;
          } else if (st == 3) {
            const int qr = L->reflect[g->wall_dim[sn]][q];
            *out = g->f[qr * nt + n];
          } else if (st == 4) {
            *out = g->f[q * nt + n];
          }
        }
      }
  double* t = g->f; g->f = g->f_buf; g->f_buf = t;
}

/* Grid::clear_forces, couplb_lattice.h:243-247 */
static void o_clear_forces(void* vg) {
  grid_t* g = (grid_t*)vg;
  memset(g->fx, 0, sizeof(double) * (size_t)g->ntotal);
  memset(g->fy, 0, sizeof(double) * (size_t)g->ntotal);
  memset(g->fz, 0, sizeof(double) * (size_t)g->ntotal);
}

/* ------------------------------------------------------------------ */
/* IBM (couplb_ibm.h:11-133)                                           */
/* ------------------------------------------------------------------ */
double clo_delta_roma(double r) {
  const double ar = fabs(r);
  if (ar <= 0.5) {
    double arg = 1.0 - 3.0*r*r;
    if (arg < 0.0) arg = 0.0;
    return (1.0 + sqrt(arg)) / 3.0;
  }
  if (ar <= 1.5) {
    double arg = -2.0 + 6.0*ar - 3.0*r*r;
    if (arg < 0.0) arg = 0.0;
    return (5.0 - 3.0*ar - sqrt(arg)) / 6.0;
  }
  return 0.0;
}

double clo_delta_peskin4(double r) {
  const double ar = fabs(r);
  if (ar <= 1.0) {
    double arg = 1.0 + 4.0*ar - 4.0*ar*ar;
    if (arg < 0.0) arg = 0.0;
    return (3.0 - 2.0*ar + sqrt(arg)) / 8.0;
  }
  if (ar <= 2.0) {
    double arg = -7.0 + 12.0*ar - 4.0*ar*ar;
    if (arg < 0.0) arg = 0.0;
    return (5.0 - 2.0*ar - sqrt(arg)) / 8.0;
  }
  return 0.0;
}

static double o_delta(int kern, double r) {
  return kern == CLO_ROMA ? clo_delta_roma(r) : clo_delta_peskin4(r);
}

static inline double delta_product(int D, int kern, double rx, double ry, double rz) {
  return o_delta(kern, rx) * o_delta(kern, ry) * ((D == 3) ? o_delta(kern, rz) : 1.0);
}

/* IBM::phys_to_local with offset 0 (one rank), couplb_ibm.h:49-55 */
static inline void to_local(const grid_t* g, double xp, double yp, double zp,
                            const double lo[3], double inv_dx,
                            double* lx, double* ly, double* lz) {
  *lx = (xp - lo[0]) * inv_dx - 0 + 1.0;
  *ly = (yp - lo[1]) * inv_dx - 0 + 1.0;
  *lz = (g->D == 3) ? ((zp - lo[2]) * inv_dx - 0 + 1.0) : 0.0;
}

/* IBM::interpolate, couplb_ibm.h:57-90 */
static void o_interpolate(void* vg, double xp, double yp, double zp,
                          const double lo[3], double inv_dx, int kern, double out[3]) {
  const grid_t* g = (const grid_t*)vg;
  double lx, ly, lz;
  to_local(g, xp, yp, zp, lo, inv_dx, &lx, &ly, &lz);
  const int i0 = (int)floor(lx), j0 = (int)floor(ly);
  const int k0 = (g->D == 3) ? (int)floor(lz) : 0;
  double ui = 0, vi = 0, wi = 0, ws = 0;
  const int klo = (g->D == 3) ? k0 - 1 : 0, khi = (g->D == 3) ? k0 + 2 : 0;
  for (int k = klo; k <= khi; k++) {
    if (g->D == 3 && (k < 0 || k >= g->gz)) continue;
    for (int j = j0 - 1; j <= j0 + 2; j++) {
      if (j < 0 || j >= g->gy) continue;
      for (int i = i0 - 1; i <= i0 + 2; i++) {
        if (i < 0 || i >= g->gx) continue;
        const long n = gidx(g, i, j, k);
        const int nt = g->type[n];
        const double d = delta_product(g->D, kern, lx - i, ly - j, lz - k);
        if (nt == 0) { ui += g->ux[n]*d; vi += g->uy[n]*d; wi += g->uz[n]*d; ws += d; }
        else if (nt == 1) { ws += d; }
        else if (nt == 2) { ui += g->bcx[n]*d; vi += g->bcy[n]*d; wi += g->bcz[n]*d; ws += d; }
        else if (nt == 3) { ws += d; }
      }
    }
  }
  if (ws > DELTA_TOL) { const double iw = 1.0 / ws; out[0] = ui*iw; out[1] = vi*iw; out[2] = wi*iw; }
  else { out[0] = out[1] = out[2] = 0.0; }
}

/* IBM::spread, couplb_ibm.h:92-133 */
static void o_spread(void* vg, double xp, double yp, double zp,
                     double fxl, double fyl, double fzl, double dvl,
                     const double lo[3], double inv_dx, double dx_phys, int kern) {
  grid_t* g = (grid_t*)vg;
  double lx, ly, lz;
  to_local(g, xp, yp, zp, lo, inv_dx, &lx, &ly, &lz);
  const int i0 = (int)floor(lx), j0 = (int)floor(ly);
  const int k0 = (g->D == 3) ? (int)floor(lz) : 0;
  const double cell_vol = dx_phys * dx_phys * ((g->D == 3) ? dx_phys : 1.0);
  const double dvr = dvl / cell_vol;
  double ws = 0.0, dc[64];
  long nc[64];
  int cnt = 0;
  const int klo = (g->D == 3) ? k0 - 1 : 0, khi = (g->D == 3) ? k0 + 2 : 0;
  for (int k = klo; k <= khi; k++) {
    if (g->D == 3 && (k < 0 || k >= g->gz)) continue;
    for (int j = j0 - 1; j <= j0 + 2; j++) {
      if (j < 0 || j >= g->gy) continue;
      for (int i = i0 - 1; i <= i0 + 2; i++) {
        if (i < 0 || i >= g->gx) continue;
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        const double d = delta_product(g->D, kern, lx - i, ly - j, lz - k);
        dc[cnt] = d; nc[cnt] = n; cnt++;
        ws += d;
      }
    }
  }
  if (ws < DELTA_TOL) return;
  const double sc = dvr / ws;
  for (int c = 0; c < cnt; c++) {
    const double w = dc[c] * sc;
    g->fx[nc[c]] += fxl * w;
    g->fy[nc[c]] += fyl * w;
    g->fz[nc[c]] += fzl * w;
  }
}

/* Grid::max_velocity / compute_diagnostics, couplb_lattice.h:203-242 */
static double o_max_velocity(void* vg) {
  grid_t* g = (grid_t*)vg;
  double umax2 = 0.0;
  const int klo = (g->D == 3) ? 1 : 0, khi = (g->D == 3) ? g->gz - 2 : 0;
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= g->gy - 2; j++)
      for (int i = 1; i <= g->gx - 2; i++) {
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        const double u2 = g->ux[n]*g->ux[n] + g->uy[n]*g->uy[n] + g->uz[n]*g->uz[n];
        if (u2 > umax2) umax2 = u2;
      }
  return sqrt(umax2);
}
static void o_diagnostics(void* vg, double out[4]) {
  grid_t* g = (grid_t*)vg;
  double m = 0, px = 0, py = 0, pz = 0;
  const int klo = (g->D == 3) ? 1 : 0, khi = (g->D == 3) ? g->gz - 2 : 0;
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= g->gy - 2; j++)
      for (int i = 1; i <= g->gx - 2; i++) {
        const long n = gidx(g, i, j, k);
        if (g->type[n] != 0) continue;
        m += g->rho[n];
        px += g->rho[n] * g->ux[n];
        py += g->rho[n] * g->uy[n];
        pz += g->rho[n] * g->uz[n];
      }
  out[0] = m; out[1] = px; out[2] = py; out[3] = pz;
}

static const clo_ops ORACLE_OPS = {
  "oracle-c",
  o_create, o_destroy, o_dfield, o_ifield, o_dims,
  o_macro, o_collide, o_exchange, o_stream, o_exchange_velocity,
  o_exchange_forces, o_clear_forces, o_interpolate, o_spread,
  o_max_velocity, o_diagnostics, o_delta
};

const clo_ops* clo_oracle_ops(void) { return &ORACLE_OPS; }

/* ------------------------------------------------------------------ */
/* Fix layer: FixCoupLB orchestration (fix_couplb.cpp)                 */
/* ------------------------------------------------------------------ */
struct clo_fix {
  const clo_ops* ops;
  void* g;
  clo_fix_params p;
  int D;
  double dx, inv_dx, dt_LBM, vel_scale, nu_lb, tau, force_scale;
  double gext[3];
  int md_sub_count;
  int ibm_has_particles;
  long next_check, next_output;
  long lbm_step_count;
  double last_diag[6];
  double* prof_rho; double* prof_ux; double* prof_uy;
  int have_profile;
  int gx, gy, gz; long ntotal; int Q;
};

static void seterr(char* err, int errlen, const char* msg) {
  if (err && errlen > 0) { strncpy(err, msg, (size_t)errlen - 1); err[errlen - 1] = 0; }
}

/* FixCoupLB::FixCoupLB validation (fix_couplb.cpp:71-238) and ::init
 * unit conversion and checks (fix_couplb.cpp:256-349), then setup_grid and
 * setup_boundaries (484-539). */
clo_fix* clo_fix_create(const clo_ops* ops, const clo_fix_params* p, char* err, int errlen) {
  if (p->Nx <= 0 || p->Ny <= 0 || p->Nz <= 0) { seterr(err, errlen, "fix couplb: Nx, Ny, and Nz must be positive"); return NULL; }
  if (p->nu <= 0.0 || !isfinite(p->nu)) { seterr(err, errlen, "fix couplb: viscosity must be positive and finite"); return NULL; }
  if (p->rho <= 0.0 || !isfinite(p->rho)) { seterr(err, errlen, "fix couplb: density must be positive and finite"); return NULL; }
  if (p->md_per_lb < 1) { seterr(err, errlen, "fix couplb: md_per_lb must be >= 1"); return NULL; }
  if (p->xi_ibm <= 0.0 || p->xi_ibm > 1.0 || !isfinite(p->xi_ibm)) { seterr(err, errlen, "fix couplb: xi_ibm must satisfy 0 < xi_ibm <= 1"); return NULL; }
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++)
      if (p->wall[d][s] < 0 || p->wall[d][s] > 4) { seterr(err, errlen, "fix couplb: wall types must be in [0,4]"); return NULL; }
  if (p->dim == 2 && p->Nz != 1) { seterr(err, errlen, "fix couplb: Nz must be 1 for 2D"); return NULL; }

  clo_fix* fx = (clo_fix*)calloc(1, sizeof(clo_fix));
  fx->ops = ops; fx->p = *p; fx->D = p->dim;
  const int is3d = (p->dim == 3);
  const double dx_x = (p->boxhi[0] - p->boxlo[0]) / p->Nx;
  const double dx_y = (p->boxhi[1] - p->boxlo[1]) / p->Ny;
  if (fabs(dx_x - dx_y) > 1e-10 * dx_x) { seterr(err, errlen, "fix couplb: non-cubic grid cells"); free(fx); return NULL; }
  if (is3d) {
    const double dx_z = (p->boxhi[2] - p->boxlo[2]) / p->Nz;
    if (fabs(dx_x - dx_z) > 1e-10 * dx_x) { seterr(err, errlen, "fix couplb: non-cubic grid cells"); free(fx); return NULL; }
  }
  fx->dx = dx_x;
  fx->inv_dx = 1.0 / fx->dx;
  fx->dt_LBM = p->dt * p->md_per_lb;
  fx->vel_scale = fx->dx / fx->dt_LBM;
  fx->nu_lb = p->nu * fx->dt_LBM / (fx->dx * fx->dx);
  fx->tau = fx->nu_lb / (1.0/3.0) + 0.5;
  if (fx->tau <= 0.5) { seterr(err, errlen, "fix couplb: tau <= 0.5 (unstable)"); free(fx); return NULL; }
  const int dim = is3d ? 3 : 2;
  double dp = 1.0;
  for (int d = 0; d < dim + 1; d++) dp *= fx->dx;
  fx->force_scale = p->rho * dp / (fx->dt_LBM * fx->dt_LBM);
  if (fx->force_scale <= 0.0 || !isfinite(fx->force_scale)) { seterr(err, errlen, "fix couplb: force_scale is non-positive or non-finite"); free(fx); return NULL; }
  const double inv_vel = 1.0 / fx->vel_scale;
  double wv[3] = { p->wall_vel[0]*inv_vel, p->wall_vel[1]*inv_vel, p->wall_vel[2]*inv_vel };
  const double cs = sqrt(1.0/3.0);
  const double uw = sqrt(wv[0]*wv[0] + wv[1]*wv[1] + wv[2]*wv[2]);
  if (uw > MA_ERROR*cs) { seterr(err, errlen, "fix couplb: wall Ma > 0.5"); free(fx); return NULL; }
  if (p->grav_on[0] && p->grav_on[1] && p->grav_on[2]) {
    const double ac = fx->dt_LBM * fx->dt_LBM / fx->dx;
    const double a = p->g[0]*ac, b = p->g[1]*ac, c = p->g[2]*ac;
    const double gm = sqrt(a*a + b*b + c*c);
    if (gm > ZERO_TOL) {
      const int H = (p->wall[1][0] > 0 && p->wall[1][1] > 0) ? p->Ny :
                    (p->wall[2][0] > 0 && p->wall[2][1] > 0) ? p->Nz : 0;
      if (H > 0) {
        const double ue = gm*H*H/(8.0*fx->nu_lb);
        if (ue/cs > MA_ERROR) { seterr(err, errlen, "fix couplb: estimated Poiseuille Ma > 0.5"); free(fx); return NULL; }
      }
    }
  }
  for (int d = 0; d < 3; d++) fx->gext[d] = p->g[d];
  fx->next_output = (p->output_every > 0) ? p->ntimestep + p->output_every : 0;
  fx->next_check = (p->check_every > 0) ? p->ntimestep + p->check_every : -1;

  clo_grid_params gp;
  memset(&gp, 0, sizeof(gp));
  gp.D = dim; gp.nx = p->Nx; gp.ny = p->Ny; gp.nz = is3d ? p->Nz : 1;
  gp.tau = fx->tau; gp.rho_lb = 1.0;
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      gp.wall[d][s] = (d < dim) ? p->wall[d][s] : 0;
      for (int c = 0; c < 3; c++)
        gp.face_vel[d][s][c] = p->use_face_vel ? p->face_vel[d][s][c] * inv_vel : wv[c];
    }
  fx->g = ops->create(&gp);
  int gd[3];
  ops->dims(fx->g, gd, &fx->ntotal, &fx->Q);
  fx->gx = gd[0]; fx->gy = gd[1]; fx->gz = gd[2];
  fx->prof_rho = dalloc(p->Ny, 0.0); fx->prof_ux = dalloc(p->Ny, 0.0); fx->prof_uy = dalloc(p->Ny, 0.0);
  return fx;
}

void clo_fix_destroy(clo_fix* fx) {
  if (!fx) return;
  fx->ops->destroy(fx->g);
  free(fx->prof_rho); free(fx->prof_ux); free(fx->prof_uy);
  free(fx);
}

void* clo_fix_grid(clo_fix* fx) { return fx->g; }
void clo_fix_set_gravity(clo_fix* fx, double gx, double gy, double gz) {
  fx->gext[0] = gx; fx->gext[1] = gy; fx->gext[2] = gz;
}

double clo_fix_scalar(clo_fix* fx, const char* name) {
  if (!strcmp(name, "tau")) return fx->tau;
  if (!strcmp(name, "nu_lb")) return fx->nu_lb;
  if (!strcmp(name, "dx")) return fx->dx;
  if (!strcmp(name, "dt_LBM")) return fx->dt_LBM;
  if (!strcmp(name, "vel_scale")) return fx->vel_scale;
  if (!strcmp(name, "force_scale")) return fx->force_scale;
  if (!strcmp(name, "lbm_steps")) return (double)fx->lbm_step_count;
  if (!strcmp(name, "md_sub_count")) return (double)fx->md_sub_count;
  return NAN;
}

/* apply_external_force, fix_couplb.cpp:680-708 (constant-style gravity; the
 * variable style re-evaluates gext before this call). */
static void apply_external_force(clo_fix* fx) {
  const clo_fix_params* p = &fx->p;
  if (!p->grav_on[0] && !p->grav_on[1] && !p->grav_on[2]) return;
  const double ac = fx->dt_LBM * fx->dt_LBM / fx->dx;
  const double gx = fx->gext[0] * ac, gy = fx->gext[1] * ac, gz = fx->gext[2] * ac;
  if (fabs(gx) < ZERO_TOL && fabs(gy) < ZERO_TOL && fabs(gz) < ZERO_TOL) return;
  double* rho = fx->ops->dfield(fx->g, CLO_F_RHO);
  double* fxa = fx->ops->dfield(fx->g, CLO_F_FX);
  double* fya = fx->ops->dfield(fx->g, CLO_F_FY);
  double* fza = fx->ops->dfield(fx->g, CLO_F_FZ);
  const int* type = fx->ops->ifield(fx->g, CLO_I_TYPE);
  const int klo = (fx->D == 3) ? 1 : 0, khi = (fx->D == 3) ? fx->gz - 2 : 0;
  const int gxn = fx->gx, gyn = fx->gy;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = klo; k <= khi; k++)
    for (int j = 1; j <= gyn - 2; j++)
      for (int i = 1; i <= gxn - 2; i++) {
        const long n = (long)i + (long)gxn * ((long)j + (long)gyn * k);
        if (type[n] != 0) continue;
        fxa[n] += gx * rho[n]; fya[n] += gy * rho[n]; fza[n] += gz * rho[n];
      }
}

/* enforce_wall_ghost_fields, fix_couplb.cpp:800-809 (serial, all nodes) */
static void enforce_wall_ghost_fields(clo_fix* fx) {
  double* rho = fx->ops->dfield(fx->g, CLO_F_RHO);
  double* ux = fx->ops->dfield(fx->g, CLO_F_UX);
  double* uy = fx->ops->dfield(fx->g, CLO_F_UY);
  double* uz = fx->ops->dfield(fx->g, CLO_F_UZ);
  const double* bx = fx->ops->dfield(fx->g, CLO_F_BCUX);
  const double* by = fx->ops->dfield(fx->g, CLO_F_BCUY);
  const double* bz = fx->ops->dfield(fx->g, CLO_F_BCUZ);
  const int* type = fx->ops->ifield(fx->g, CLO_I_TYPE);
  for (long n = 0; n < fx->ntotal; n++) {
    if (type[n] == 1 || type[n] == 3) { rho[n] = 1.0; ux[n] = uy[n] = uz[n] = 0; }
    else if (type[n] == 2) { rho[n] = 1.0; ux[n] = bx[n]; uy[n] = by[n]; uz[n] = bz[n]; }
  }
}

/* do_lbm_step, fix_couplb.cpp:659-669 */
static void do_lbm_step(clo_fix* fx) {
  apply_external_force(fx);
  fx->ops->compute_macroscopic(fx->g, 1);
  fx->ops->collide(fx->g);
  fx->ops->exchange(fx->g);
  fx->ops->stream(fx->g);
  enforce_wall_ghost_fields(fx);
  fx->ops->clear_forces(fx->g);
}

/* do_ibm_sub_coupling, fix_couplb.cpp:730-798 */
static void do_ibm_sub_coupling(clo_fix* fx, const clo_atoms* a, int first) {
  int ng = 0;
  for (int i = 0; i < a->nlocal; i++) if (a->mask[i] & a->groupbit) ng++;
  if (ng == 0) { fx->ibm_has_particles = 0; return; }
  if (first || !fx->ibm_has_particles) {
    fx->ops->compute_macroscopic(fx->g, 0);
    fx->ops->exchange_velocity(fx->g);
  }
  fx->ibm_has_particles = 1;
  const double xi_over_dtLBM = fx->p.xi_ibm / fx->dt_LBM;
  const double inv_N = 1.0 / (double)fx->p.md_per_lb;
  const double dv = (fx->D == 3) ? (fx->dx*fx->dx*fx->dx) : (fx->dx*fx->dx);
  for (int i = 0; i < a->nlocal; i++) {
    if (!(a->mask[i] & a->groupbit)) continue;
    const double* x = &a->x[3*i];
    const double* v = &a->v[3*i];
    double uf[3];
    fx->ops->interpolate(fx->g, x[0], x[1], x[2], fx->p.boxlo, fx->inv_dx, fx->p.kernel, uf);
    const double m = a->mass[i];
    const double fxi = m * xi_over_dtLBM * (uf[0]*fx->vel_scale - v[0]);
    const double fyi = m * xi_over_dtLBM * (uf[1]*fx->vel_scale - v[1]);
    const double fzi = m * xi_over_dtLBM * (uf[2]*fx->vel_scale - v[2]);
    a->f[3*i+0] += fxi; a->f[3*i+1] += fyi; a->f[3*i+2] += fzi;
    fx->ops->spread(fx->g, x[0], x[1], x[2],
                    -fxi * inv_N / fx->force_scale,
                    -fyi * inv_N / fx->force_scale,
                    -fzi * inv_N / fx->force_scale,
                    dv, fx->p.boxlo, fx->inv_dx, fx->dx, fx->p.kernel);
  }
  if (fx->md_sub_count == fx->p.md_per_lb - 1) fx->ops->exchange_forces(fx->g);
}

/* write_profile's y-binned averages, fix_couplb.cpp:903-979 (one rank) */
static void compute_profile(clo_fix* fx, long step) {
  const double* rho = fx->ops->dfield(fx->g, CLO_F_RHO);
  const double* ux = fx->ops->dfield(fx->g, CLO_F_UX);
  const double* uy = fx->ops->dfield(fx->g, CLO_F_UY);
  const int nx = fx->p.Nx, ny = fx->p.Ny, nz = (fx->D == 3) ? fx->p.Nz : 1;
  FILE* fp = NULL;
  if (fx->p.output_file) fp = fopen(fx->p.output_file, "a");
  if (fp) fprintf(fp, "# step = %ld\n", step);
  for (int j = 0; j < ny; j++) {
    double su = 0, sv = 0, sr = 0;
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nx; i++) {
        const long n = (long)(i + 1) + (long)fx->gx * ((long)(j + 1) + (long)fx->gy * ((fx->D == 3) ? k + 1 : 0));
        su += ux[n]; sv += uy[n]; sr += rho[n];
      }
    const int cnt = nx * nz;
    const double inv = (cnt > 0) ? 1.0 / cnt : 0.0;
    fx->prof_rho[j] = fx->p.rho * (sr * inv);
    fx->prof_ux[j] = fx->vel_scale * (su * inv);
    fx->prof_uy[j] = fx->vel_scale * (sv * inv);
    if (fp) fprintf(fp, "%ld %d %.6f %.8f %.10e %.10e\n", step, j,
                    fx->p.boxlo[1] + (double)j * fx->dx,
                    fx->prof_rho[j], fx->prof_ux[j], fx->prof_uy[j]);
  }
  if (fp) { fprintf(fp, "\n"); fclose(fp); }
  fx->have_profile = 1;
}

static void write_profile_header(clo_fix* fx) {
  if (!fx->p.output_file) return;
  const char* mode = (fx->p.ntimestep == 0) ? "w" : "a";
  FILE* fp = fopen(fx->p.output_file, mode);
  if (!fp) return;
  if (fx->p.ntimestep == 0) {
    fprintf(fp, "# H=Ny=%d tau=%.6f nu=%.4e nu_lb=%.6f rho=%.4f kernel=%s\n",
            fx->p.Ny, fx->tau, fx->p.nu, fx->nu_lb, fx->p.rho,
            fx->p.kernel == CLO_ROMA ? "roma" : "peskin4");
    fprintf(fp, "# CoupLB velocity profile\n# Columns: step j y rho ux uy\n#\n");
  }
  fclose(fp);
}

/* FixCoupLB::setup, fix_couplb.cpp:551-638 */
void clo_fix_setup(clo_fix* fx, const clo_atoms* a) {
  fx->md_sub_count = 0;
  do_ibm_sub_coupling(fx, a, 1);
  if (fx->p.md_per_lb > 1) fx->ops->exchange_forces(fx->g);
  fx->md_sub_count = 0;
  if (fx->p.output_every > 0) {
    fx->ops->compute_macroscopic(fx->g, 0);
    write_profile_header(fx);
    compute_profile(fx, fx->p.ntimestep);
  }
}

/* FixCoupLB::post_force, fix_couplb.cpp:640-651 */
void clo_fix_post_force(clo_fix* fx, const clo_atoms* a) {
  const int first = (fx->md_sub_count == 0);
  if (first) { do_lbm_step(fx); fx->lbm_step_count++; }
  do_ibm_sub_coupling(fx, a, first);
  fx->md_sub_count = (fx->md_sub_count + 1) % fx->p.md_per_lb;
}

/* FixCoupLB::end_of_step + check_stability_precomputed, fix_couplb.cpp:818-901 */
int clo_fix_end_of_step(clo_fix* fx, long step) {
  int chk = 0, out = 0;
  if (fx->p.check_every > 0 && step >= fx->next_check) { chk = 1; fx->next_check = step + fx->p.check_every; }
  if (fx->p.output_every > 0 && step >= fx->next_output) { out = 1; chk = 1; fx->next_output += fx->p.output_every; }
  if (!chk && !out) return 0;
  const int clamps = fx->ops->compute_macroscopic(fx->g, 0);
  int rc = 0;
  if (chk) {
    const double umax = fx->ops->max_velocity(fx->g);
    const double ma = umax / sqrt(1.0/3.0);
    double d[4];
    fx->ops->diagnostics(fx->g, d);
    fx->last_diag[0] = ma; fx->last_diag[1] = d[0]; fx->last_diag[2] = d[1];
    fx->last_diag[3] = d[2]; fx->last_diag[4] = d[3];
    fx->last_diag[5] = (double)clamps;
    if (ma > MA_ERROR) rc = 2;
  }
  if (out) compute_profile(fx, step);
  return rc;
}

void clo_fix_last_diag(clo_fix* fx, double out[6]) { memcpy(out, fx->last_diag, sizeof(double) * 6); }

int clo_fix_last_profile(clo_fix* fx, double* rho, double* ux, double* uy) {
  if (!fx->have_profile) return 0;
  memcpy(rho, fx->prof_rho, sizeof(double) * (size_t)fx->p.Ny);
  memcpy(ux, fx->prof_ux, sizeof(double) * (size_t)fx->p.Ny);
  memcpy(uy, fx->prof_uy, sizeof(double) * (size_t)fx->p.Ny);
  return fx->p.Ny;
}

/* read_checkpoint's state restore, couplb_io.h:813-817 */
void clo_fix_set_f(clo_fix* fx, const double* f) {
  double* fa = fx->ops->dfield(fx->g, CLO_F_F);
  double* fb = fx->ops->dfield(fx->g, CLO_F_FBUF);
  memcpy(fa, f, sizeof(double) * (size_t)(fx->Q * fx->ntotal));
  memcpy(fb, f, sizeof(double) * (size_t)(fx->Q * fx->ntotal));
  fx->ops->compute_macroscopic(fx->g, 0);
}

/* ------------------------------------------------------------------ */
/* Checkpoint, couplb_io.h:653-828                                     */
/* ------------------------------------------------------------------ */
#define CLBK_MAGIC 0x434C424Bu
#define CLBK_VERSION 1

int clo_write_checkpoint(const char* fname, int D, int Q, const int dims[3],
                         const int offs[3], const int glob[3], long step,
                         const double* f, long nf) {
  FILE* fp = fopen(fname, "wb");
  if (!fp) return 0;
  uint32_t magic = CLBK_MAGIC; int32_t ver = CLBK_VERSION; int64_t s64 = step;
  int32_t d32 = D, q32 = Q, a[3], b[3], c[3];
  for (int i = 0; i < 3; i++) { a[i] = dims[i]; b[i] = offs[i]; c[i] = glob[i]; }
  int ok = 1;
  ok = ok && fwrite(&magic, 4, 1, fp) == 1;
  ok = ok && fwrite(&ver, 4, 1, fp) == 1;
  ok = ok && fwrite(&s64, 8, 1, fp) == 1;
  ok = ok && fwrite(&d32, 4, 1, fp) == 1;
  ok = ok && fwrite(&q32, 4, 1, fp) == 1;
  ok = ok && fwrite(a, 4, 3, fp) == 3;
  ok = ok && fwrite(b, 4, 3, fp) == 3;
  ok = ok && fwrite(c, 4, 3, fp) == 3;
  ok = ok && fwrite(f, sizeof(double), (size_t)nf, fp) == (size_t)nf;
  fclose(fp);
  return ok;
}

int clo_read_checkpoint(const char* fname, int D, int Q, const int dims[3],
                        const int offs[3], const int glob[3], long* step,
                        double* f, long nf, char* err, int errlen) {
  FILE* fp = fopen(fname, "rb");
  if (!fp) { seterr(err, errlen, "cannot open"); return 0; }
  uint32_t magic; int32_t ver; int64_t s64; int32_t d32, q32, a[3], b[3], c[3];
  if (fread(&magic, 4, 1, fp) != 1 || magic != CLBK_MAGIC) { seterr(err, errlen, "bad magic"); fclose(fp); return 0; }
  if (fread(&ver, 4, 1, fp) != 1 || ver != CLBK_VERSION) { seterr(err, errlen, "unsupported version"); fclose(fp); return 0; }
  if (fread(&s64, 8, 1, fp) != 1) { seterr(err, errlen, "truncated header (step)"); fclose(fp); return 0; }
  if (fread(&d32, 4, 1, fp) != 1 || fread(&q32, 4, 1, fp) != 1) { seterr(err, errlen, "truncated header (D/Q)"); fclose(fp); return 0; }
  if (d32 != D || q32 != Q) { seterr(err, errlen, "D/Q mismatch"); fclose(fp); return 0; }
  if (fread(a, 4, 3, fp) != 3 || fread(b, 4, 3, fp) != 3 || fread(c, 4, 3, fp) != 3) { seterr(err, errlen, "truncated header (dims)"); fclose(fp); return 0; }
  if (a[0] != dims[0] || a[1] != dims[1] || a[2] != dims[2]) { seterr(err, errlen, "local grid mismatch"); fclose(fp); return 0; }
  if (b[0] != offs[0] || b[1] != offs[1] || b[2] != offs[2]) { seterr(err, errlen, "offset mismatch"); fclose(fp); return 0; }
  if (c[0] != glob[0] || c[1] != glob[1] || c[2] != glob[2]) { seterr(err, errlen, "global grid mismatch"); fclose(fp); return 0; }
  size_t nr = fread(f, sizeof(double), (size_t)nf, fp);
  fclose(fp);
  if (nr != (size_t)nf) { seterr(err, errlen, "short read"); return 0; }
  *step = (long)s64;
  return 1;
}

/* ------------------------------------------------------------------ */
/* CPU baseline timing: do_lbm_step's passes on a bare grid            */
/* ------------------------------------------------------------------ */
static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

double clo_time_lbm_steps(const clo_ops* ops, void* g, int steps,
                          double gx, double gy, double gz) {
  int gd[3]; long nt; int Q;
  ops->dims(g, gd, &nt, &Q);
  double* rho = ops->dfield(g, CLO_F_RHO);
  double* fxa = ops->dfield(g, CLO_F_FX);
  double* fya = ops->dfield(g, CLO_F_FY);
  double* fza = ops->dfield(g, CLO_F_FZ);
  double* ux = ops->dfield(g, CLO_F_UX);
  double* uy = ops->dfield(g, CLO_F_UY);
  double* uz = ops->dfield(g, CLO_F_UZ);
  const int* type = ops->ifield(g, CLO_I_TYPE);
  const int is3d = gd[2] > 1;
  const int klo = is3d ? 1 : 0, khi = is3d ? gd[2] - 2 : 0;
  const double t0 = now_s();
  for (int s = 0; s < steps; s++) {
    if (!(fabs(gx) < ZERO_TOL && fabs(gy) < ZERO_TOL && fabs(gz) < ZERO_TOL)) {
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = klo; k <= khi; k++)
        for (int j = 1; j <= gd[1] - 2; j++)
          for (int i = 1; i <= gd[0] - 2; i++) {
            const long n = (long)i + (long)gd[0] * ((long)j + (long)gd[1] * k);
            if (type[n] != 0) continue;
            fxa[n] += gx * rho[n]; fya[n] += gy * rho[n]; fza[n] += gz * rho[n];
          }
    }
    ops->compute_macroscopic(g, 1);
    ops->collide(g);
    ops->exchange(g);
    ops->stream(g);
    for (long n = 0; n < nt; n++) {
      if (type[n] == 1 || type[n] == 3) { rho[n] = 1.0; ux[n] = uy[n] = uz[n] = 0; }
    }
    ops->clear_forces(g);
  }
  return now_s() - t0;
}

int clo_omp_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
