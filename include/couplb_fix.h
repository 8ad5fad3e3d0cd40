/*
 * couplb_fix.h -- C ABI over the host-side FixCoupLB core (libcouplb_hip.so).
 *
 * The core (couplb_amd/host/fix_couplb_core.cpp) is the C++ host code of
 * `fix ID group couplb Nx Ny Nz nu rho0 [keywords]`: it parses the same
 * arguments as FixCoupLB::FixCoupLB (fix_couplb.cpp:71-238), performs
 * FixCoupLB::init's unit conversion and checks (256-349) and runs the
 * callbacks setup / post_force / end_of_step (551-901), calling the device
 * C ABI in couplb_hip.h.  A LAMMPS build binds it through the thin adapter
 * couplb_amd/lammps/fix_couplb.{h,cpp}; these C entry points expose the same
 * callbacks to other hosts (the Python tests, ctypes).
 *
 * Errors the reference raises with error->all / error->one return
 * COUPLB_EINVAL with the reference's message in couplb_fix_last_error().
 */
#ifndef COUPLB_FIX_H
#define COUPLB_FIX_H

#include "couplb_hip.h" /* status codes COUPLB_OK, COUPLB_EINVAL, ... */

#ifdef __cplusplus
extern "C" {
#endif

typedef struct couplb_fix couplb_fix;

/* What LAMMPS' Domain/Update hold and FixCoupLB reads
 * (fix_couplb.cpp:236, 282, 305, 486-499). */
typedef struct couplb_domain {
  int dimension;
  double boxlo[3], boxhi[3];
  int periodicity[3];
  double dt;           /* update->dt */
  long ntimestep;      /* update->ntimestep at run start */
} couplb_domain;

/* LAMMPS' per-atom arrays (atom->x, v, f as [nlocal][3], mask, rmass or
 * mass[type]); x, v read, f added into (fix_couplb.cpp:733-785). */
typedef struct couplb_atoms {
  int nlocal;
  const double* x;
  const double* v;
  double* f;
  const int* mask;
  int groupbit;
  const double* rmass;   /* per-atom mass, or NULL */
  const double* mass;    /* per-type mass (1-based like LAMMPS), used when rmass is NULL */
  const int* type;
} couplb_atoms;

/* Per-atom arrays that only `vtk_solid` reads (write_solid_vtk,
 * couplb_io.h:335-377); NULL where the atom style lacks them. */
typedef struct couplb_atom_extras {
  const void* tag;        /* atom->tag */
  int tag_bytes;          /* sizeof(tagint): 4 or 8 */
  const void* molecule;   /* atom->molecule (tagint, tag_bytes each), or NULL */
  const double* radius;   /* atom->radius, or NULL */
  const double* mu;       /* &atom->mu[0][0]: 4 values per atom, or NULL */
  const double* omega;    /* &atom->omega[0][0], or NULL */
  const double* torque;   /* &atom->torque[0][0], or NULL */
} couplb_atom_extras;

/* Equal-style variable lookup and evaluation for `gravity v_name ...`.
 * value == NULL: existence check only, as FixCoupLB::init does with
 * input->variable->find / equalstyle (fix_couplb.cpp:263-280), nothing is
 * evaluated.  Otherwise evaluate into *value (compute_equal at every LB
 * step, fix_couplb.cpp:671-678).  Return 0 on success,
 * COUPLB_VAR_MISSING for the reference's "does not exist" error,
 * COUPLB_VAR_NOT_EQUAL for its "must be equal-style" error. */
#define COUPLB_VAR_MISSING 1
#define COUPLB_VAR_NOT_EQUAL 2
typedef int (*couplb_var_eval)(void* user, const char* name, double* value);

/* arg = {"ID", "group", "couplb", "Nx", "Ny", "Nz", "nu", "rho", keywords...} */
int couplb_fix_create(int narg, const char* const* arg, const couplb_domain* dom,
                      couplb_fix** out);
int couplb_fix_destroy(couplb_fix* fx);
const char* couplb_fix_last_error(const couplb_fix* fx);
int couplb_fix_set_variable_callback(couplb_fix* fx, couplb_var_eval fn, void* user);
/* refresh box, timestep and periodicity before init (FixCoupLB::init reads
 * domain->boxlo/boxhi and update->dt, fix_couplb.cpp:282, 305) */
int couplb_fix_set_domain(couplb_fix* fx, const couplb_domain* dom);
int couplb_fix_init(couplb_fix* fx);                     /* FixCoupLB::init */
int couplb_fix_setup(couplb_fix* fx, const couplb_atoms* a);      /* ::setup */
/* ::post_force.  On a grid whose steps fuse (couplb_lbm_fuses_steps) the
 * LB step of a sub-step with no group atoms is held back and issued with
 * the following ones as one run, at the latest by the next output, check,
 * VTK or checkpoint step, IBM sub-step with atoms, setup or context
 * access, a change of gravity, or after 512 steps; results are the same. */
int couplb_fix_post_force(couplb_fix* fx, const couplb_atoms* a);
/* ::end_of_step; VTK output reads the atom arrays handed to the last
 * setup / post_force call (the same timestep's arrays in LAMMPS) */
int couplb_fix_end_of_step(couplb_fix* fx, long step);
/* the extra per-atom arrays of vtk_solid; refresh them with the atom
 * arrays before setup / post_force (they move when LAMMPS reallocates) */
int couplb_fix_set_atom_extras(couplb_fix* fx, const couplb_atom_extras* e);
int couplb_fix_setmask(void);                            /* POST_FORCE|END_OF_STEP */

/* Introspection used by tests and the bench. */
double couplb_fix_scalar(const couplb_fix* fx, const char* name);
/* last stability check: {Ma, mass, momx, momy, momz, clamps} */
int couplb_fix_last_diag(const couplb_fix* fx, double out[6]);
/* last y-profile in physical units: Ny rows each of rho, ux, uy */
int couplb_fix_last_profile(const couplb_fix* fx, double* rho, double* ux, double* uy);
/* device context, for state import/export; issues the LB steps the fix
 * still holds back first (see couplb_fix_post_force) */
void* couplb_fix_context(couplb_fix* fx);
/* captured screen output (the reference prints to LAMMPS' screen) */
const char* couplb_fix_log(const couplb_fix* fx);

#ifdef __cplusplus
}
#endif
#endif
