/*
 * couplb_oracle.h -- CPU FP64 restatement of the CoupLB LB/IBM step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the oracle for the HIP product path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in couplb_amd/ links or calls it.
 *
 * It follows the reference algorithm file by file (citations are relative to
 * the reference checkout):
 *   lattice tables ........ CoupLB/couplb_lattice.h:24-75
 *   Grid storage/macro .... CoupLB/couplb_lattice.h:82-248
 *   BGK + Guo ............. CoupLB/couplb_collision.h:18-62
 *   stream / fills / fold . CoupLB/couplb_streaming.h:82-293 (self-periodic branches)
 *   wall typing ........... CoupLB/couplb_boundary.h:12-38
 *   IBM delta/interp/spread CoupLB/couplb_ibm.h:11-133
 *   fix orchestration ..... CoupLB/fix_couplb.cpp:256-349, 484-539, 551-565,
 *                           640-809, 818-844, 846-979
 *   checkpoint ............ CoupLB/couplb_io.h:653-828
 *
 * The layout is the reference's ghost-inclusive SoA, f[q*ntotal + n] with
 * n = i + gx*(j + gy*k), so outputs compare element-for-element with a build
 * of the reference headers.
 *
 * Two layers:
 *   clo_ops  -- a table of the numerics primitives (the Grid/BGK/Streaming/
 *               IBM operations).  clo_oracle_ops() is this file's own C
 *               restatement; oracle/_ref provides the same table backed by
 *               the reference's own headers.
 *   clo_fix  -- the FixCoupLB orchestration (init/setup/post_force/
 *               end_of_step) written once on top of any clo_ops table.
 */
#ifndef COUPLB_ORACLE_H
#define COUPLB_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { CLO_ROMA = 0, CLO_PESKIN4 = 1 };

/* field selectors for clo_ops.dfield / ifield */
enum {
  CLO_F_RHO = 0, CLO_F_UX, CLO_F_UY, CLO_F_UZ,
  CLO_F_FX, CLO_F_FY, CLO_F_FZ,
  CLO_F_BCUX, CLO_F_BCUY, CLO_F_BCUZ,
  CLO_F_F, CLO_F_FBUF
};
enum { CLO_I_TYPE = 0, CLO_I_WALLDIM = 1 };

typedef struct clo_grid_params {
  int D;               /* 2 (D2Q9) or 3 (D3Q19) */
  int nx, ny, nz;      /* interior nodes (nz = 1 in 2D) */
  double tau;
  int wall[3][2];      /* face types 0..4 per dim (lo, hi) */
  double face_vel[3][2][3]; /* lattice-unit velocity of type-2 faces */
  double rho_lb;       /* reference density in lattice units (1.0) */
} clo_grid_params;

typedef struct clo_ops {
  const char* name;
  void*   (*create)(const clo_grid_params*);
  void    (*destroy)(void*);
  double* (*dfield)(void*, int which);
  int*    (*ifield)(void*, int which);
  void    (*dims)(void*, int g[3], long* ntotal, int* Q);
  int     (*compute_macroscopic)(void*, int guo); /* returns clamp count */
  void    (*collide)(void*);
  void    (*exchange)(void*);
  void    (*stream)(void*);
  void    (*exchange_velocity)(void*);
  void    (*exchange_forces)(void*);
  void    (*clear_forces)(void*);
  void    (*interpolate)(void*, double xp, double yp, double zp,
                         const double lo[3], double inv_dx, int kern,
                         double out[3]);
  void    (*spread)(void*, double xp, double yp, double zp,
                    double fxl, double fyl, double fzl, double dvl,
                    const double lo[3], double inv_dx, double dx_phys,
                    int kern);
  double  (*max_velocity)(void*);
  void    (*diagnostics)(void*, double out[4]);
  double  (*delta)(int kern, double r);
} clo_ops;

const clo_ops* clo_oracle_ops(void);

/* ------------------------------------------------------------------ */
/* Fix layer                                                           */
/* ------------------------------------------------------------------ */
typedef struct clo_fix_params {
  int dim;                 /* 2 or 3 */
  int Nx, Ny, Nz;
  double nu, rho;          /* physical viscosity and density */
  int md_per_lb;
  double xi_ibm;
  int grav_on[3];          /* 0 = keyword absent, 1 = constant */
  double g[3];             /* physical acceleration */
  int wall[3][2];
  double wall_vel[3];      /* physical velocity for every type-2 face */
  int use_face_vel;        /* extension: per-face velocities below */
  double face_vel[3][2][3];
  int kernel;
  int check_every, output_every;
  double boxlo[3], boxhi[3];
  double dt;               /* MD timestep */
  long ntimestep;          /* LAMMPS step at run start */
  const char* output_file; /* NULL: profile kept in memory only */
} clo_fix_params;

typedef struct clo_atoms {
  int nlocal;
  const double* x;   /* [nlocal][3] */
  const double* v;   /* [nlocal][3] */
  double* f;         /* [nlocal][3], the fix adds into it */
  const int* mask;
  int groupbit;
  const double* mass; /* per-atom mass (rmass, or mass[type]) */
} clo_atoms;

typedef struct clo_fix clo_fix;

/* Returns NULL and fills err on an init-time error (the reference's
 * error->all at fix_couplb.cpp:71-238, 256-349). */
clo_fix* clo_fix_create(const clo_ops* ops, const clo_fix_params* p,
                        char* err, int errlen);
void  clo_fix_destroy(clo_fix*);
void  clo_fix_setup(clo_fix*, const clo_atoms*);
void  clo_fix_post_force(clo_fix*, const clo_atoms*);
/* returns 0, or 2 when the Ma>0.5 error fires (fix_couplb.cpp:823) */
int   clo_fix_end_of_step(clo_fix*, long step);
void  clo_fix_set_gravity(clo_fix*, double gx, double gy, double gz);
void* clo_fix_grid(clo_fix*);
double clo_fix_scalar(clo_fix*, const char* name);
/* last diagnostics: Ma, mass, momx, momy, momz, clamps */
void  clo_fix_last_diag(clo_fix*, double out[6]);
/* last y-profile: rho, ux, uy per row (physical units), Ny rows */
int   clo_fix_last_profile(clo_fix*, double* rho, double* ux, double* uy);
/* ghost-inclusive f import (restart path) */
void  clo_fix_set_f(clo_fix*, const double* f);

/* Checkpoint (.clbk) in the reference's format, couplb_io.h:653-828.
 * Return 1 on success, 0 on failure (message in err). */
int clo_write_checkpoint(const char* fname, int D, int Q, const int dims[3],
                         const int offs[3], const int glob[3], long step,
                         const double* f, long nf);
int clo_read_checkpoint(const char* fname, int D, int Q, const int dims[3],
                        const int offs[3], const int glob[3], long* step,
                        double* f, long nf, char* err, int errlen);

/* Standalone helpers for the delta kernels (couplb_ibm.h:11-33). */
double clo_delta_roma(double r);
double clo_delta_peskin4(double r);

/* Timed CPU baseline: runs `steps` LB steps of a grid in the reference's
 * pass structure (do_lbm_step, fix_couplb.cpp:659-669) and returns wall
 * seconds.  Used by the bench's cpu_baseline leg. */
double clo_time_lbm_steps(const clo_ops* ops, void* grid, int steps,
                          double gx, double gy, double gz);
int clo_omp_threads(void);

#ifdef __cplusplus
}
#endif
#endif
