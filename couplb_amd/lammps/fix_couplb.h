// fix_couplb.h -- LAMMPS adapter for the MI355X build of `fix couplb`.
//
// Drop-in replacement for CoupLB/fix_couplb.h (same style name, same
// callbacks, fix_couplb.h:1-30).  The class owns a couplb_fix handle
// (include/couplb_fix.h) and forwards each callback to it; all grid work
// runs on the GPU through include/couplb_hip.h.  Compiled only inside a
// LAMMPS source tree (it includes LAMMPS headers).
// clang-format off
#ifdef FIX_CLASS
FixStyle(couplb,FixCoupLB);
#else

#ifndef LMP_FIX_COUPLB_MI355X_H
#define LMP_FIX_COUPLB_MI355X_H

#include "fix.h"
#include "couplb_fix.h"

namespace LAMMPS_NS {

class FixCoupLB : public Fix {
 public:
  FixCoupLB(class LAMMPS *, int, char **);
  ~FixCoupLB() override;
  int setmask() override;
  void init() override;
  void setup(int) override;
  void post_force(int) override;
  void end_of_step() override;

 private:
  couplb_fix *core;
  couplb_domain current_domain() const;
  couplb_atoms atom_view() const;
  void send_extras();
  void check(int rc);
  static int eval_variable(void *self, const char *name, double *value);
};

}  // namespace LAMMPS_NS

#endif
#endif
