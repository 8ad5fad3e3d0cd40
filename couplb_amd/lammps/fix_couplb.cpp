// fix_couplb.cpp -- LAMMPS adapter for the MI355X build of `fix couplb`.
// Each callback forwards to the host core through include/couplb_fix.h; the
// core reproduces FixCoupLB's behaviour (CoupLB/fix_couplb.cpp).
#include "fix_couplb.h"

#include "atom.h"
#include "comm.h"
#include "domain.h"
#include "error.h"
#include "input.h"
#include "modify.h"
#include "update.h"
#include "variable.h"

using namespace LAMMPS_NS;
using namespace FixConst;

FixCoupLB::FixCoupLB(LAMMPS *lmp, int narg, char **arg) : Fix(lmp, narg, arg), core(nullptr)
{
  // the whole grid lives on one GPU: one MPI rank drives it
  if (comm->nprocs != 1)
    error->all(FLERR, "fix couplb (MI355X build) runs on one MPI rank holding the whole grid");
  couplb_domain dom = current_domain();
  const int rc = couplb_fix_create(narg, arg, &dom, &core);
  if (rc != COUPLB_OK) {
    const std::string msg = couplb_fix_last_error(core);
    couplb_fix_destroy(core);
    core = nullptr;
    error->all(FLERR, msg);
  }
  couplb_fix_set_variable_callback(core, &FixCoupLB::eval_variable, this);
}

FixCoupLB::~FixCoupLB() { couplb_fix_destroy(core); }

int FixCoupLB::setmask() { return couplb_fix_setmask() == 3 ? (POST_FORCE | END_OF_STEP) : 0; }

couplb_domain FixCoupLB::current_domain() const
{
  couplb_domain d;
  d.dimension = domain->dimension;
  for (int k = 0; k < 3; k++) {
    d.boxlo[k] = domain->boxlo[k];
    d.boxhi[k] = domain->boxhi[k];
    d.periodicity[k] = domain->periodicity[k];
  }
  d.dt = update->dt;
  d.ntimestep = (long) update->ntimestep;
  return d;
}

couplb_atoms FixCoupLB::atom_view() const
{
  couplb_atoms a;
  a.nlocal = atom->nlocal;
  a.x = atom->nlocal ? &atom->x[0][0] : nullptr;
  a.v = atom->nlocal ? &atom->v[0][0] : nullptr;
  a.f = atom->nlocal ? &atom->f[0][0] : nullptr;
  a.mask = atom->mask;
  a.groupbit = groupbit;
  a.rmass = atom->rmass;
  a.mass = atom->mass;
  a.type = atom->type;
  return a;
}

void FixCoupLB::check(int rc)
{
  if (rc != COUPLB_OK) error->all(FLERR, couplb_fix_last_error(core));
}

// input->variable lookup for `gravity v_name` (fix_couplb.cpp:263-280, 671-678).
// The variables are evaluated again at the next LB step, md_per_lb MD steps
// later, so the computes they use are scheduled for that step
// (addstep_compute(ntimestep + md_per_lb), fix_couplb.cpp:677); repeating
// clearstep / addstep per variable schedules the same step.
int FixCoupLB::eval_variable(void *self, const char *name, double *value)
{
  auto *fx = static_cast<FixCoupLB *>(self);
  const int ivar = fx->input->variable->find(name);
  if (ivar < 0) return COUPLB_VAR_MISSING;
  if (!fx->input->variable->equalstyle(ivar)) return COUPLB_VAR_NOT_EQUAL;
  if (!value) return 0;  // init: existence check only
  fx->modify->clearstep_compute();
  *value = fx->input->variable->compute_equal(ivar);
  const long md_per_lb = (long) couplb_fix_scalar(fx->core, "md_per_lb");
  fx->modify->addstep_compute(fx->update->ntimestep + md_per_lb);
  return 0;
}

void FixCoupLB::init()
{
  couplb_domain dom = current_domain();
  check(couplb_fix_set_domain(core, &dom));
  check(couplb_fix_init(core));
  if (comm->me == 0 && screen) fputs(couplb_fix_log(core), screen);
}

// per-atom arrays only vtk_solid reads (couplb_io.h:335-377)
void FixCoupLB::send_extras()
{
  couplb_atom_extras e;
  const int n = atom->nlocal;
  e.tag = atom->tag;
  e.tag_bytes = (int) sizeof(tagint);
  e.molecule = atom->molecule;
  e.radius = atom->radius;
  e.mu = (atom->mu && n) ? &atom->mu[0][0] : nullptr;
  e.omega = (atom->omega && n) ? &atom->omega[0][0] : nullptr;
  e.torque = (atom->torque && n) ? &atom->torque[0][0] : nullptr;
  check(couplb_fix_set_atom_extras(core, &e));
}

void FixCoupLB::setup(int)
{
  couplb_atoms a = atom_view();
  send_extras();
  check(couplb_fix_setup(core, &a));
}

void FixCoupLB::post_force(int)
{
  couplb_atoms a = atom_view();
  send_extras();
  check(couplb_fix_post_force(core, &a));
}

void FixCoupLB::end_of_step()
{
  check(couplb_fix_end_of_step(core, (long) update->ntimestep));
}
