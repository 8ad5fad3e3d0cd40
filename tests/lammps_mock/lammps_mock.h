// Test-only stand-in for the few LAMMPS classes and members that the
// adapter couplb_amd/lammps/fix_couplb.cpp touches, declared with LAMMPS'
// public names and types, so the adapter compiles and runs outside a LAMMPS
// tree.  Error::all throws instead of aborting the MPI job.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#define FLERR __FILE__, __LINE__

namespace LAMMPS_NS {

typedef int tagint;

namespace FixConst {
enum { POST_FORCE = 1 << 7, END_OF_STEP = 1 << 9 };
}

class Atom {
 public:
  int nlocal = 0;
  double **x = nullptr, **v = nullptr, **f = nullptr;
  int *mask = nullptr, *type = nullptr;
  double *rmass = nullptr, *mass = nullptr, *radius = nullptr;
  tagint *tag = nullptr, *molecule = nullptr;
  double **mu = nullptr, **omega = nullptr, **torque = nullptr;
};

class Comm {
 public:
  int me = 0, nprocs = 1;
};

class Domain {
 public:
  int dimension = 3;
  double boxlo[3] = {0, 0, 0}, boxhi[3] = {1, 1, 1};
  int periodicity[3] = {1, 1, 1};
};

class Error {
 public:
  [[noreturn]] void all(const std::string &file, int line, const std::string &msg)
  {
    throw std::runtime_error(file + ":" + std::to_string(line) + ": " + msg);
  }
};

// one equal-style variable, gx = 1e-5 * (1 + 0.01 * step)
class Variable {
 public:
  const long *ntimestep = nullptr;
  int find(const char *name) { return std::string(name) == "gx" ? 0 : -1; }
  int equalstyle(int ivar) { return ivar == 0; }
  double compute_equal(int) { return 1e-5 * (1.0 + 0.01 * (ntimestep ? (double) *ntimestep : 0.0)); }
};

class Input {
 public:
  Variable *variable = nullptr;
};

// records the steps the computes are scheduled for
class Modify {
 public:
  std::vector<long> addsteps;
  void clearstep_compute() {}
  void addstep_compute(long step) { addsteps.push_back(step); }
};

class Update {
 public:
  double dt = 1.0;
  long ntimestep = 0;
};

class LAMMPS {
 public:
  Atom *atom = nullptr;
  Comm *comm = nullptr;
  Domain *domain = nullptr;
  Error *error = nullptr;
  Input *input = nullptr;
  Modify *modify = nullptr;
  Update *update = nullptr;
  FILE *screen = nullptr;
};

class Fix {
 public:
  Fix(LAMMPS *lmp, int, char **)
      : atom(lmp->atom), comm(lmp->comm), domain(lmp->domain), error(lmp->error),
        input(lmp->input), modify(lmp->modify), update(lmp->update), screen(lmp->screen)
  {
  }
  virtual ~Fix() = default;
  virtual int setmask() = 0;
  virtual void init() {}
  virtual void setup(int) {}
  virtual void post_force(int) {}
  virtual void end_of_step() {}
  int groupbit = 1;

 protected:
  Atom *atom;
  Comm *comm;
  Domain *domain;
  Error *error;
  Input *input;
  Modify *modify;
  Update *update;
  FILE *screen;
};

}  // namespace LAMMPS_NS
