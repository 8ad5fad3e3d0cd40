// fix_couplb_core.hpp -- host-side FixCoupLB on the MI355X device context.
//
// The class keeps the reference fix's interface and behaviour
// (CoupLB/fix_couplb.h:22-30, fix_couplb.cpp) -- same argument syntax,
// validation messages, unit conversion, callback order and sub-stepping --
// while every grid operation goes through the C ABI in couplb_hip.h.
#pragma once
#include <cstddef>
#include <cstdio>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/couplb_fix.h"
#include "../../include/couplb_hip.h"

namespace couplb_host {

// A double array in pinned host memory from the device library
// (couplb_host_alloc); grows, never shrinks.
class PinnedVec {
 public:
  PinnedVec() = default;
  ~PinnedVec() { couplb_host_free(p_); }
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  void resize(size_t n) {
    if (n > cap_) {
      couplb_host_free(p_);
      p_ = static_cast<double*>(couplb_host_alloc(n * sizeof(double)));
      cap_ = p_ ? n : 0;
      if (!p_) throw std::bad_alloc();
    }
    n_ = n;
  }
  void assign(size_t n, double v) {
    resize(n);
    for (size_t i = 0; i < n; i++) p_[i] = v;
  }
  double* data() { return p_; }
  const double* data() const { return p_; }
  size_t size() const { return n_; }
  double& operator[](size_t i) { return p_[i]; }
  double operator[](size_t i) const { return p_[i]; }

 private:
  double* p_ = nullptr;
  size_t n_ = 0, cap_ = 0;
};

// error->all / error->one (LAMMPS aborts; here the C boundary reports it)
struct fix_error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class FixCoupLBCore {
 public:
  FixCoupLBCore(int narg, const char* const* arg, const couplb_domain& dom);
  ~FixCoupLBCore();
  static int setmask();
  void init();
  void setup(const couplb_atoms& a);
  void post_force(const couplb_atoms& a);
  int end_of_step(long step);

  void set_variable_callback(couplb_var_eval fn, void* user) { var_fn = fn; var_user = user; }
  void set_atom_extras(const couplb_atom_extras& e) { extras = e; have_extras = true; }
  void set_domain(const couplb_domain& d) { dom = d; }
  double scalar(const std::string& name) const;
  const double* last_diag() const { return diag; }
  bool last_profile(double* r, double* u, double* v) const;
  // the device context with every held-back LB step issued
  couplb_ctx* context();
  const std::string& log() const { return screen; }

 private:
  enum { NONE_G, CONSTANT_G, EQUAL_G };
  couplb_domain dom;
  int Nx, Ny, Nz;
  double nu, rho;
  int md_per_lb = 1, md_sub_count = 0;
  double xi_ibm = 1.0;
  double gext[3] = {0, 0, 0};
  int gstyle[3] = {NONE_G, NONE_G, NONE_G};
  std::string gvar[3];
  int wall[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  double wall_vel[3] = {0, 0, 0};
  bool face_vel_set[3][2] = {{false, false}, {false, false}, {false, false}};
  double face_vel[3][2][3] = {};
  int output_every = 0;
  std::string output_file = "couplb_profile.dat";
  int check_every = 0;
  int vtk_every = 0;
  std::string vtk_prefix = "couplb_vtk", vtk_pvd_file;
  bool has_vtk_region = false;
  double vtk_rlo[3] = {0, 0, 0}, vtk_rhi[3] = {0, 0, 0};
  int vtk_region[6] = {0, 0, 0, 0, 0, 0};
  bool vtk_solid_on = false;
  std::vector<std::string> vtk_solid_attrs;
  std::string vtk_solid_pvd_file;
  std::vector<long> vtk_steps;
  int checkpoint_every = 0;
  std::string checkpoint_prefix = "couplb_ckpt";
  std::string restart_prefix;
  bool do_restart = false;
  int kernel = COUPLB_ROMA;
  int storage = COUPLB_FP64;
  double storage_range = 0.0;                 // COUPLB_I16 range (0: library default)
  unsigned long long i16_saturations_seen = 0;  // counter value at the last check
  bool is3d = true;

  double dx = 0, inv_dx = 0, dt_LBM = 0, vel_scale = 0, nu_lb = 0, tau = 0, force_scale = 0;
  long next_output = 0, next_check = -1, next_checkpoint = -1, next_vtk = -1;
  long lbm_step_count = 0;
  bool ibm_has_particles = false;
  couplb_ctx* ctx = nullptr;
  // LB steps of a small grid with no atoms in the group are held back and
  // issued as one run (couplb_lbm_steps_async fuses them into one launch)
  bool defer_steps = false;
  int pend_n = 0, pend_flags = 0;
  double pend_g[3] = {0, 0, 0};
  couplb_var_eval var_fn = nullptr;
  void* var_user = nullptr;

  double diag[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> prof_r, prof_u, prof_v;
  bool have_profile = false;
  std::string screen;

  // packed group atoms in pinned host memory (the sub-step's copies to and
  // from the device run as DMA)
  bool masses_sent = false;  // pm has been sent to the device by a host sub-step
  PinnedVec px, pv, pm, pf;
  std::vector<double> px_prev;  // positions of the previous sub-step (reuse test)
  std::vector<int> pidx;
  couplb_atoms atoms_now{};  // the arrays of the current step (setup / post_force)
  bool have_atoms = false;
  couplb_atom_extras extras{};
  bool have_extras = false;

  void say(const char* fmt, ...);
  void check(int rc, const char* what);
  void update_gravity_variables();
  void lbm_step(int ng);
  int flush_steps();
  int count_group(const couplb_atoms& a) const;
  void ibm_sub_coupling(const couplb_atoms& a, bool first);
  void check_stability_precomputed(long step);
  void write_profile(long step);
  void write_checkpoint(long step);
  void write_vtk(long step);
  void write_solid_vtk(long step);
  void read_checkpoint();
  bool refresh_follows_next_step() const;
};

}  // namespace couplb_host
