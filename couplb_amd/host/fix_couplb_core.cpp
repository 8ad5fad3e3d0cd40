// fix_couplb_core.cpp -- host-side FixCoupLB on the MI355X device context.
// Behaviour follows CoupLB/fix_couplb.cpp; citations give the line ranges.
#include "fix_couplb_core.hpp"

#include <algorithm>
#include <functional>
#include <set>

#include "vtk_io.hpp"

#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <string>

namespace couplb_host {

namespace {

constexpr double ZERO_TOL = 1e-15;  // couplb_lattice.h:16
constexpr double MA_WARN = 0.3;     // couplb_lattice.h:18
constexpr double MA_ERROR = 0.5;    // couplb_lattice.h:19
// held-back LB steps issued as one run at the latest after this many
constexpr int MAX_HELD_STEPS = 512;
constexpr uint32_t CLBK_MAGIC = 0x434C424Bu;  // couplb_io.h:19-23
constexpr int32_t CLBK_VERSION = 1;

// keyword whitelist, fix_couplb.cpp:22-36, plus the extensions
// `storage` (population precision), `storage_range` (range of the 16-bit
// fixed point) and `wall_vel_face` (per-face wall velocity)
bool is_keyword(const char* s) {
  static const char* kw[] = {"md_per_lb", "xi_ibm", "gravity", "wall_x", "wall_y", "wall_z",
                             "wall_vel", "output", "check_every", "kernel", "vtk", "vtk_region",
                             "vtk_solid", "checkpoint", "restart", "storage", "storage_range",
                             "wall_vel_face", nullptr};
  for (int i = 0; kw[i]; i++)
    if (std::strcmp(s, kw[i]) == 0) return true;
  return false;
}

int to_int(const char* s) {
  try {
    return std::stoi(s);
  } catch (...) {
    throw fix_error(std::string("Expected integer parameter instead of '") + s + "'");
  }
}
double to_double(const char* s) {
  try {
    return std::stod(s);
  } catch (...) {
    throw fix_error(std::string("Expected floating point parameter instead of '") + s + "'");
  }
}

}  // namespace

void FixCoupLBCore::say(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  screen += buf;
}

void FixCoupLBCore::check(int rc, const char* what) {
  if (rc != COUPLB_OK)
    throw fix_error(std::string("fix couplb: ") + what + ": " + couplb_last_error(ctx));
}

// FixCoupLB::FixCoupLB, fix_couplb.cpp:71-238
FixCoupLBCore::FixCoupLBCore(int narg, const char* const* arg, const couplb_domain& d) : dom(d) {
  if (narg < 8) throw fix_error("Illegal fix couplb command: need Nx Ny Nz nu rho");
  Nx = to_int(arg[3]);
  Ny = to_int(arg[4]);
  Nz = to_int(arg[5]);
  nu = to_double(arg[6]);
  rho = to_double(arg[7]);
  if (Nx <= 0 || Ny <= 0 || Nz <= 0) throw fix_error("fix couplb: Nx, Ny, and Nz must be positive");
  if (nu <= 0.0 || !std::isfinite(nu)) throw fix_error("fix couplb: viscosity must be positive and finite");
  if (rho <= 0.0 || !std::isfinite(rho)) throw fix_error("fix couplb: density must be positive and finite");
  int iarg = 8;
  while (iarg < narg) {
    const char* k = arg[iarg];
    if (!is_keyword(k)) throw fix_error(std::string("Unknown fix couplb keyword: ") + k);
    auto need = [&](int n, const char* msg) {
      if (iarg + n + 1 > narg) throw fix_error(msg);
    };
    if (!std::strcmp(k, "md_per_lb")) {
      need(1, "fix couplb md_per_lb: need 1 value");
      md_per_lb = to_int(arg[iarg + 1]);
      if (md_per_lb < 1) throw fix_error("fix couplb: md_per_lb must be >= 1");
      iarg += 2;
    } else if (!std::strcmp(k, "xi_ibm")) {
      need(1, "fix couplb xi_ibm: need 1 value");
      xi_ibm = to_double(arg[iarg + 1]);
      if (xi_ibm <= 0.0 || xi_ibm > 1.0 || !std::isfinite(xi_ibm))
        throw fix_error("fix couplb: xi_ibm must satisfy 0 < xi_ibm <= 1");
      iarg += 2;
    } else if (!std::strcmp(k, "gravity")) {
      need(3, "fix couplb gravity: need 3 values");
      for (int c = 0; c < 3; c++) {
        const char* a = arg[iarg + 1 + c];
        if (std::strncmp(a, "v_", 2) == 0) {
          gvar[c] = a + 2;
          gstyle[c] = EQUAL_G;
        } else {
          gext[c] = to_double(a);
          gstyle[c] = CONSTANT_G;
        }
      }
      iarg += 4;
    } else if (!std::strcmp(k, "wall_x") || !std::strcmp(k, "wall_y") || !std::strcmp(k, "wall_z")) {
      const int dd = k[5] - 'x';
      const char msg[3][40] = {"fix couplb wall_x: need 2 values", "fix couplb wall_y: need 2 values",
                               "fix couplb wall_z: need 2 values"};
      need(2, msg[dd]);
      wall[dd][0] = to_int(arg[iarg + 1]);
      wall[dd][1] = to_int(arg[iarg + 2]);
      if (wall[dd][0] < 0 || wall[dd][0] > 4 || wall[dd][1] < 0 || wall[dd][1] > 4)
        throw fix_error(std::string("fix couplb: ") + k + " types must be in [0,4]");
      iarg += 3;
    } else if (!std::strcmp(k, "wall_vel")) {
      need(3, "fix couplb wall_vel: need 3 values");
      for (int c = 0; c < 3; c++) wall_vel[c] = to_double(arg[iarg + 1 + c]);
      iarg += 4;
    } else if (!std::strcmp(k, "wall_vel_face")) {
      // wall_vel_face x|y|z lo|hi vx vy vz  (per-face velocity, extension)
      need(5, "fix couplb wall_vel_face: need dim side vx vy vz");
      const char* dn = arg[iarg + 1];
      const char* sd = arg[iarg + 2];
      const int dd = (!std::strcmp(dn, "x")) ? 0 : (!std::strcmp(dn, "y")) ? 1 : (!std::strcmp(dn, "z")) ? 2 : -1;
      const int ss = (!std::strcmp(sd, "lo")) ? 0 : (!std::strcmp(sd, "hi")) ? 1 : -1;
      if (dd < 0 || ss < 0) throw fix_error("fix couplb wall_vel_face: use x|y|z lo|hi");
      for (int c = 0; c < 3; c++) face_vel[dd][ss][c] = to_double(arg[iarg + 3 + c]);
      face_vel_set[dd][ss] = true;
      iarg += 6;
    } else if (!std::strcmp(k, "output")) {
      need(2, "fix couplb output: need 2 values");
      output_every = to_int(arg[iarg + 1]);
      output_file = arg[iarg + 2];
      if (output_every < 0) throw fix_error("fix couplb: output frequency cannot be negative");
      iarg += 3;
    } else if (!std::strcmp(k, "check_every")) {
      need(1, "fix couplb check_every: need 1 value");
      check_every = to_int(arg[iarg + 1]);
      if (check_every < 0) throw fix_error("fix couplb: check_every cannot be negative");
      iarg += 2;
    } else if (!std::strcmp(k, "kernel")) {
      need(1, "fix couplb kernel: need 1 value");
      if (!std::strcmp(arg[iarg + 1], "roma")) kernel = COUPLB_ROMA;
      else if (!std::strcmp(arg[iarg + 1], "peskin4")) kernel = COUPLB_PESKIN4;
      else throw fix_error("fix couplb kernel: use roma or peskin4");
      iarg += 2;
    } else if (!std::strcmp(k, "storage")) {
      need(1, "fix couplb storage: need 1 value");
      const char* s = arg[iarg + 1];
      if (!std::strcmp(s, "fp64")) storage = COUPLB_FP64;
      else if (!std::strcmp(s, "fp32")) storage = COUPLB_FP32;
      else if (!std::strcmp(s, "fp16")) storage = COUPLB_FP16;
      else if (!std::strcmp(s, "bf16")) storage = COUPLB_BF16;
      else if (!std::strcmp(s, "i16")) storage = COUPLB_I16;
      else throw fix_error("fix couplb storage: use fp64, fp32, fp16, bf16 or i16");
      iarg += 2;
    } else if (!std::strcmp(k, "storage_range")) {
      need(1, "fix couplb storage_range: need 1 value");
      storage_range = to_double(arg[iarg + 1]);
      if (!(storage_range > 0.0)) throw fix_error("fix couplb storage_range must be > 0");
      iarg += 2;
    } else if (!std::strcmp(k, "vtk")) {
      need(2, "fix couplb vtk: need N prefix");
      vtk_every = to_int(arg[iarg + 1]);
      vtk_prefix = arg[iarg + 2];
      if (vtk_every < 0) throw fix_error("fix couplb: vtk frequency cannot be negative");
      vtk_pvd_file = vtk_prefix + ".pvd";
      iarg += 3;
    } else if (!std::strcmp(k, "vtk_region")) {
      iarg++;
      int nreg = 0;
      while (iarg + nreg < narg && !is_keyword(arg[iarg + nreg])) nreg++;
      if (nreg != 4 && nreg != 6)
        throw fix_error("fix couplb vtk_region: need 4 values (2D: xlo xhi ylo yhi) or 6 values (3D: xlo xhi ylo yhi zlo zhi)");
      for (int d = 0; d < 2; d++) {
        vtk_rlo[d] = to_double(arg[iarg + 2 * d]);
        vtk_rhi[d] = to_double(arg[iarg + 2 * d + 1]);
      }
      vtk_rlo[2] = (nreg == 6) ? to_double(arg[iarg + 4]) : -1e30;
      vtk_rhi[2] = (nreg == 6) ? to_double(arg[iarg + 5]) : 1e30;
      has_vtk_region = true;
      iarg += nreg;
    } else if (!std::strcmp(k, "vtk_solid")) {
      if (vtk_solid_on) throw fix_error("fix couplb vtk_solid: keyword specified more than once");
      vtk_solid_on = true;
      iarg++;
      while (iarg < narg && !is_keyword(arg[iarg])) vtk_solid_attrs.push_back(arg[iarg++]);
      if (vtk_solid_attrs.empty()) throw fix_error("fix couplb vtk_solid: need at least one attribute");
    } else if (!std::strcmp(k, "checkpoint")) {
      need(2, "fix couplb checkpoint: need N prefix");
      checkpoint_every = to_int(arg[iarg + 1]);
      checkpoint_prefix = arg[iarg + 2];
      if (checkpoint_every < 0) throw fix_error("fix couplb: checkpoint frequency cannot be negative");
      iarg += 3;
    } else if (!std::strcmp(k, "restart")) {
      need(1, "fix couplb restart: need prefix");
      restart_prefix = arg[iarg + 1];
      do_restart = true;
      iarg += 2;
    }
  }
  is3d = (dom.dimension == 3);
  if (!is3d && Nz != 1) throw fix_error("fix couplb: Nz must be 1 for 2D");
}

FixCoupLBCore::~FixCoupLBCore() {
  if (ctx) couplb_destroy(ctx);
}

int FixCoupLBCore::setmask() { return 1 | 2; }  // POST_FORCE | END_OF_STEP

// FixCoupLB::init, fix_couplb.cpp:256-482 (grid and boundaries: 484-539)
void FixCoupLBCore::init() {
  if (md_per_lb < 1) throw fix_error("fix couplb: md_per_lb must be >= 1");
  if (xi_ibm <= 0.0 || xi_ibm > 1.0 || !std::isfinite(xi_ibm))
    throw fix_error("fix couplb: xi_ibm must satisfy 0 < xi_ibm <= 1");
  static const char* axis[3] = {"x", "y", "z"};
  for (int c = 0; c < 3; c++) {
    if (gstyle[c] != EQUAL_G) continue;
    const int rc = var_fn ? var_fn(var_user, gvar[c].c_str(), nullptr) : COUPLB_VAR_MISSING;
    if (rc == COUPLB_VAR_NOT_EQUAL)
      throw fix_error("Variable " + gvar[c] + " for gravity " + axis[c] + " in fix couplb must be equal-style");
    if (rc != 0)
      throw fix_error("Variable " + gvar[c] + " for gravity " + axis[c] + " in fix couplb does not exist");
  }
  const double dx_x = (dom.boxhi[0] - dom.boxlo[0]) / Nx;
  const double dx_y = (dom.boxhi[1] - dom.boxlo[1]) / Ny;
  char buf[256];
  if (std::fabs(dx_x - dx_y) > 1e-10 * dx_x) {
    std::snprintf(buf, sizeof(buf), "fix couplb: non-cubic grid cells (Lx/Nx=%.6e != Ly/Ny=%.6e). Adjust Nx, Ny, or domain size.", dx_x, dx_y);
    throw fix_error(buf);
  }
  if (is3d) {
    const double dx_z = (dom.boxhi[2] - dom.boxlo[2]) / Nz;
    if (std::fabs(dx_x - dx_z) > 1e-10 * dx_x) {
      std::snprintf(buf, sizeof(buf), "fix couplb: non-cubic grid cells (Lx/Nx=%.6e != Lz/Nz=%.6e). Adjust Nx, Nz, or domain size.", dx_x, dx_z);
      throw fix_error(buf);
    }
  }
  dx = dx_x;
  inv_dx = 1.0 / dx;
  dt_LBM = dom.dt * md_per_lb;
  vel_scale = dx / dt_LBM;
  nu_lb = nu * dt_LBM / (dx * dx);
  tau = nu_lb / (1.0 / 3.0) + 0.5;
  if (tau <= 0.5) throw fix_error("fix couplb: tau <= 0.5 (unstable)");
  if (tau < 0.505) say("WARNING: fix couplb: tau very close to 0.5\n");
  const int dim = is3d ? 3 : 2;
  double dp = 1.0;
  for (int d = 0; d < dim + 1; d++) dp *= dx;
  force_scale = rho * dp / (dt_LBM * dt_LBM);
  if (force_scale <= 0.0 || !std::isfinite(force_scale))
    throw fix_error("fix couplb: force_scale is non-positive or non-finite (check rho, dx, dt_LBM)");
  const double inv_vel = 1.0 / vel_scale;
  const double wv[3] = {wall_vel[0] * inv_vel, wall_vel[1] * inv_vel, wall_vel[2] * inv_vel};
  const double cs = std::sqrt(1.0 / 3.0);
  double uw = std::sqrt(wv[0] * wv[0] + wv[1] * wv[1] + wv[2] * wv[2]);
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++)
      if (face_vel_set[d][s]) {
        const double* f = face_vel[d][s];
        uw = std::max(uw, std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]) * inv_vel);
      }
  if (uw > MA_ERROR * cs) throw fix_error("fix couplb: wall Ma > 0.5");
  if (uw > MA_WARN * cs) say("CoupLB WARNING: wall Ma=%.3f > 0.3\n", uw / cs);
  if (gstyle[0] == CONSTANT_G && gstyle[1] == CONSTANT_G && gstyle[2] == CONSTANT_G) {
    const double ac = dt_LBM * dt_LBM / dx;
    const double a = gext[0] * ac, b = gext[1] * ac, c = gext[2] * ac;
    const double gm = std::sqrt(a * a + b * b + c * c);
    if (gm > ZERO_TOL) {
      const int H = (wall[1][0] > 0 && wall[1][1] > 0) ? Ny : (wall[2][0] > 0 && wall[2][1] > 0) ? Nz : 0;
      if (H > 0) {
        const double ue = gm * H * H / (8.0 * nu_lb);
        if (ue / cs > MA_ERROR) throw fix_error("fix couplb: estimated Poiseuille Ma > 0.5");
        if (ue / cs > MA_WARN) say("CoupLB WARNING: estimated Poiseuille Ma=%.3f\n", ue / cs);
      }
    }
  }
  if (output_every > 0) next_output = dom.ntimestep + output_every;
  next_check = (check_every > 0) ? dom.ntimestep + check_every : -1;
  next_checkpoint = (checkpoint_every > 0) ? dom.ntimestep + checkpoint_every : -1;
  next_vtk = (vtk_every > 0) ? dom.ntimestep + vtk_every : -1;
  lbm_step_count = 0;
  md_sub_count = 0;

  // vtk_solid / vtk_region set-up, fix_couplb.cpp:378-450
  if (vtk_solid_on) {
    if (vtk_every <= 0) throw fix_error("fix couplb vtk_solid requires vtk keyword to set output frequency");
    vtk_solid_pvd_file = vtk_prefix + "_solid.pvd";
    say("CoupLB: vtk_solid with %d attributes\n", (int)vtk_solid_attrs.size());
  }
  if (has_vtk_region) {
    const int Nn[3] = {Nx, Ny, is3d ? Nz : 1};
    for (int d = 0; d < 3; d++) {
      vtk_rlo[d] = std::max(vtk_rlo[d], dom.boxlo[d]);  // clamp to the box first
      vtk_rhi[d] = std::min(vtk_rhi[d], dom.boxhi[d]);
    }
    for (int d = 0; d < 3; d++) {
      if (d == 2 && !is3d) {
        vtk_region[4] = 0;
        vtk_region[5] = 1;
        continue;
      }
      vtk_region[2 * d] = (int)std::floor((vtk_rlo[d] - dom.boxlo[d]) / dx);
      vtk_region[2 * d + 1] = (int)std::ceil((vtk_rhi[d] - dom.boxlo[d]) / dx);
      vtk_region[2 * d] = std::max(0, std::min(vtk_region[2 * d], Nn[d]));
      vtk_region[2 * d + 1] = std::max(0, std::min(vtk_region[2 * d + 1], Nn[d]));
    }
    if (vtk_region[1] <= vtk_region[0] || vtk_region[3] <= vtk_region[2] || vtk_region[5] <= vtk_region[4]) {
      char b[256];
      std::snprintf(b, sizeof(b),
                    "fix couplb vtk_region: empty or inverted region ([%d,%d]x[%d,%d]x[%d,%d]). Check that "
                    "xlo<xhi, ylo<yhi%s",
                    vtk_region[0], vtk_region[1], vtk_region[2], vtk_region[3], vtk_region[4], vtk_region[5],
                    is3d ? ", zlo<zhi" : "");
      throw fix_error(b);
    }
    for (int d = 0; d < 3; d++) {  // snap to node boundaries (the atom filter of vtk_solid)
      vtk_rlo[d] = dom.boxlo[d] + vtk_region[2 * d] * dx;
      vtk_rhi[d] = dom.boxlo[d] + vtk_region[2 * d + 1] * dx;
    }
    if (is3d)
      say("CoupLB: vtk_region [%.4f,%.4f]x[%.4f,%.4f]x[%.4f,%.4f] -> nodes [%d,%d]x[%d,%d]x[%d,%d]\n",
          vtk_rlo[0], vtk_rhi[0], vtk_rlo[1], vtk_rhi[1], vtk_rlo[2], vtk_rhi[2], vtk_region[0], vtk_region[1],
          vtk_region[2], vtk_region[3], vtk_region[4], vtk_region[5]);
    else
      say("CoupLB: vtk_region [%.4f,%.4f]x[%.4f,%.4f] -> nodes [%d,%d]x[%d,%d]\n", vtk_rlo[0], vtk_rhi[0],
          vtk_rlo[1], vtk_rhi[1], vtk_region[0], vtk_region[1], vtk_region[2], vtk_region[3]);
  }
  // setup_grid + setup_boundaries (fix_couplb.cpp:484-539): one device
  // grid holds the whole domain.
  for (int d = 0; d < dim; d++)
    if ((wall[d][0] == 0 || wall[d][1] == 0) && !dom.periodicity[d])
      say("CoupLB WARNING: %s faces without walls are treated as periodic\n", axis[d]);
  couplb_params p;
  std::memset(&p, 0, sizeof(p));
  p.dim = dim;
  p.nx = Nx; p.ny = Ny; p.nz = is3d ? Nz : 1;
  p.tau = tau;
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      p.wall[d][s] = (d < dim) ? wall[d][s] : 0;
      for (int c = 0; c < 3; c++)
        p.face_vel[d][s][c] = face_vel_set[d][s] ? face_vel[d][s][c] * inv_vel : (c == 2 && !is3d ? 0.0 : wv[c]);
    }
  p.kernel = kernel;
  p.storage = storage;
  p.storage_range = storage_range;
  p.rho0 = 1.0;  // rho_lb, fix_couplb.h:41
  p.md_per_lb = md_per_lb;
  p.device = 0;
  if (ctx) { couplb_destroy(ctx); ctx = nullptr; }
  const int rc = couplb_create(&p, &ctx);
  if (rc != COUPLB_OK) {
    const std::string msg = ctx ? couplb_last_error(ctx) : "couplb_create failed";
    if (ctx) { couplb_destroy(ctx); ctx = nullptr; }
    throw fix_error("fix couplb: device context: " + msg);
  }
  defer_steps = couplb_lbm_fuses_steps(ctx) != 0;
  pend_n = 0;
  prof_r.assign(Ny, 0.0); prof_u.assign(Ny, 0.0); prof_v.assign(Ny, 0.0);

  if (do_restart) {
    read_checkpoint();
    do_restart = false;
  }
  if (vtk_every > 0)
    say("CoupLB: VTK output every %d -> %s_*.vti (series: %s)\n", vtk_every, vtk_prefix.c_str(),
        vtk_pvd_file.c_str());
  say("CoupLB: %dD grid %dx%dx%d tau=%.6f nu=%.6f rho=%.4f\n", dim, Nx, Ny, Nz, tau, nu, rho);
  say("CoupLB: nu=%.4e -> nu_lb=%.6f\n", nu, nu_lb);
  say("CoupLB: dx=%.4e dt_LBM=%.4e vel_scale=%.4e force_scale=%.4e\n", dx, dt_LBM, vel_scale, force_scale);
  say("CoupLB: kernel=%s xi_ibm=%.4f\n", kernel == COUPLB_ROMA ? "roma" : "peskin4", xi_ibm);
  say("CoupLB: device %s storage=%s\n", couplb_version(),
      storage == COUPLB_FP64 ? "fp64" : storage == COUPLB_FP32 ? "fp32" : storage == COUPLB_FP16 ? "fp16"
      : storage == COUPLB_BF16 ? "bf16" : "i16");
}

// update_gravity_variables, fix_couplb.cpp:671-678
void FixCoupLBCore::update_gravity_variables() {
  for (int c = 0; c < 3; c++) {
    if (gstyle[c] != EQUAL_G) continue;
    double v = 0;
    if (!var_fn || var_fn(var_user, gvar[c].c_str(), &v) != 0)
      throw fix_error("Variable " + gvar[c] + " for gravity in fix couplb does not exist");
    gext[c] = v;
  }
}

int FixCoupLBCore::count_group(const couplb_atoms& a) const {
  int ng = 0;
  for (int i = 0; i < a.nlocal; i++)
    if (a.mask[i] & a.groupbit) ng++;
  return ng;
}

// lbm_step / do_lbm_step with apply_external_force, fix_couplb.cpp:653-708
void FixCoupLBCore::lbm_step(int ng) {
  double g[3] = {0, 0, 0};
  if (gstyle[0] == EQUAL_G || gstyle[1] == EQUAL_G || gstyle[2] == EQUAL_G) update_gravity_variables();
  if (!(gstyle[0] == NONE_G && gstyle[1] == NONE_G && gstyle[2] == NONE_G)) {
    const double ac = dt_LBM * dt_LBM / dx;
    for (int c = 0; c < 3; c++) g[c] = gext[c] * ac;
  }
  // With group atoms present the IBM sub-step that follows recomputes the
  // density (fix_couplb.cpp:757-760), so the step need not keep it.
  const int flags = (ng > 0) ? COUPLB_STEP_RHO_REFRESHED_NEXT : 0;
  if (ng == 0 && defer_steps) {
    // the step changes nothing the host reads until the next output, IBM
    // sub-step or context access, each of which issues the held-back run
    if (pend_n > 0 && (std::memcmp(g, pend_g, sizeof(g)) != 0 || flags != pend_flags))
      check(flush_steps(), "LB steps");
    if (pend_n == 0) {
      std::memcpy(pend_g, g, sizeof(g));
      pend_flags = flags;
    }
    if (++pend_n >= MAX_HELD_STEPS) check(flush_steps(), "LB steps");
    return;
  }
  check(flush_steps(), "LB steps");
  check(couplb_lbm_step(ctx, g, flags), "LB step");
}

int FixCoupLBCore::flush_steps() {
  if (pend_n == 0) return COUPLB_OK;
  const int n = pend_n;
  pend_n = 0;
  return couplb_lbm_steps_async(ctx, pend_g, pend_flags, n);
}

couplb_ctx* FixCoupLBCore::context() {
  if (ctx && flush_steps() != COUPLB_OK) say("CoupLB WARNING: %s\n", couplb_last_error(ctx));
  return ctx;
}

// FixCoupLB::post_force, fix_couplb.cpp:640-651
void FixCoupLBCore::post_force(const couplb_atoms& a) {
  atoms_now = a;
  have_atoms = true;
  const bool first = (md_sub_count == 0);
  if (first) {
    lbm_step(count_group(a));
    lbm_step_count++;
  }
  ibm_sub_coupling(a, first);
  md_sub_count = (md_sub_count + 1) % md_per_lb;
}

// do_ibm_sub_coupling, fix_couplb.cpp:730-798
void FixCoupLBCore::ibm_sub_coupling(const couplb_atoms& a, bool first) {
  (void)first;
  const int ng = count_group(a);
  if (ng == 0) {
    ibm_has_particles = false;
    return;
  }
  check(flush_steps(), "LB steps");
  ibm_has_particles = true;
  // the group packed in order; f goes out as zeros and the returned F is
  // added into the atoms' forces below (f += F, fix_couplb.cpp:783-785)
  const bool same_n = pm.size() == (size_t)ng;
  px.resize(3 * (size_t)ng); pv.resize(3 * (size_t)ng); pm.resize(ng); pf.assign(3 * (size_t)ng, 0.0);
  pidx.resize(ng);
  bool masses_same = same_n;
  int c = 0;
  for (int i = 0; i < a.nlocal; i++) {
    if (!(a.mask[i] & a.groupbit)) continue;
    pidx[c] = i;
    for (int d = 0; d < 3; d++) {
      px[3 * c + d] = a.x[3 * i + d];
      pv[3 * c + d] = a.v[3 * i + d];
    }
    const double mi = a.rmass ? a.rmass[i] : a.mass[a.type[i]];
    if (masses_same && std::memcmp(&pm[c], &mi, sizeof(double)) != 0) masses_same = false;
    pm[c] = mi;
    c++;
  }
  couplb_ibm_args ia;
  for (int d = 0; d < 3; d++) ia.lo[d] = dom.boxlo[d];
  ia.inv_dx = inv_dx;
  ia.xi_over_dtLB = xi_ibm / dt_LBM;
  ia.vel_scale = vel_scale;
  ia.inv_N = 1.0 / static_cast<double>(md_per_lb);
  ia.force_scale = force_scale;
  // same atoms at bit-identical positions as the previous sub-step: the
  // device keeps its marker ordering, stencils and fluid velocity
  ia.positions_unchanged =
      (px_prev.size() == px.size() && std::memcmp(px_prev.data(), px.data(), sizeof(double) * px.size()) == 0)
          ? 1 : 0;
  px_prev.assign(px.data(), px.data() + px.size());
  ia.masses_unchanged = (masses_same && masses_sent) ? 1 : 0;
  ia.forces_zero = 1;
  check(couplb_ibm_substep(ctx, ng, px.data(), pv.data(), pm.data(), pf.data(), &ia), "IBM sub-step");
  masses_sent = true;
  for (int k = 0; k < ng; k++)
    for (int d = 0; d < 3; d++) a.f[3 * pidx[k] + d] += pf[3 * k + d];
}

// FixCoupLB::setup, fix_couplb.cpp:551-638
void FixCoupLBCore::setup(const couplb_atoms& a) {
  check(flush_steps(), "LB steps");
  atoms_now = a;
  have_atoms = true;
  md_sub_count = 0;
  ibm_sub_coupling(a, true);
  md_sub_count = 0;
  if (output_every > 0 || vtk_every > 0) check(couplb_refresh_macro(ctx), "refresh");
  if (output_every > 0) {
    FILE* fp = std::fopen(output_file.c_str(), dom.ntimestep == 0 ? "w" : "a");
    if (fp) {
      if (dom.ntimestep == 0) {
        std::fprintf(fp, "# H=Ny=%d tau=%.6f nu=%.4e nu_lb=%.6f rho=%.4f kernel=%s\n", Ny, tau, nu, nu_lb,
                     rho, kernel == COUPLB_ROMA ? "roma" : "peskin4");
        std::fprintf(fp, "# CoupLB velocity profile\n# Columns: step j y rho ux uy\n#\n");
      }
      std::fclose(fp);
    }
    write_profile(dom.ntimestep);
  }
  if (vtk_every > 0) write_vtk(dom.ntimestep);
}

// check_stability_precomputed, fix_couplb.cpp:818-844
void FixCoupLBCore::check_stability_precomputed(long step) {
  double out[6];
  check(couplb_diagnostics(ctx, out), "diagnostics");
  const double cs = std::sqrt(1.0 / 3.0), ma = out[4] / cs;
  diag[0] = ma; diag[1] = out[0]; diag[2] = out[1]; diag[3] = out[2]; diag[4] = out[3]; diag[5] = out[5];
  if (ma > MA_ERROR) throw fix_error("fix couplb: Ma>0.5, simulation invalid");
  if (ma > MA_WARN) say("CoupLB WARNING: Ma=%.4f at step %ld\n", ma, step);
  say("CoupLB step %ld: Ma=%.4f mass=%.8e mom=(%.4e,%.4e,%.4e)\n", step, ma, out[0], out[1], out[2], out[3]);
  if (out[5] > 0) say("CoupLB WARNING: %d nodes had density clamped at step %ld\n", (int)out[5], step);
  if (storage == COUPLB_I16) {
    unsigned long long sat = 0;
    check(couplb_storage_saturations(ctx, &sat), "storage saturations");
    if (sat > i16_saturations_seen)
      say("CoupLB WARNING: %llu populations exceeded storage_range by step %ld\n",
          sat - i16_saturations_seen, step);
    i16_saturations_seen = sat;
  }
}

// write_profile, fix_couplb.cpp:903-979 (one device grid: no gathers)
void FixCoupLBCore::write_profile(long step) {
  std::vector<double> rs(Ny), us(Ny), vs(Ny);
  check(couplb_profile_y(ctx, rs.data(), us.data(), vs.data()), "profile");
  const int cnt = Nx * (is3d ? Nz : 1);
  const double inv = cnt > 0 ? 1.0 / cnt : 0.0;
  FILE* fp = std::fopen(output_file.c_str(), "a");
  if (!fp) say("CoupLB WARNING: cannot open %s\n", output_file.c_str());
  if (fp) std::fprintf(fp, "# step = %ld\n", step);
  for (int j = 0; j < Ny; j++) {
    prof_r[j] = rho * (rs[j] * inv);
    prof_u[j] = vel_scale * (us[j] * inv);
    prof_v[j] = vel_scale * (vs[j] * inv);
    if (fp)
      std::fprintf(fp, "%ld %d %.6f %.8f %.10e %.10e\n", step, j, dom.boxlo[1] + (double)j * dx, prof_r[j],
                   prof_u[j], prof_v[j]);
  }
  if (fp) { std::fprintf(fp, "\n"); std::fclose(fp); }
  have_profile = true;
}

// IO::write_checkpoint, couplb_io.h:669-719 (rank 0 file name prefix.0.clbk)
void FixCoupLBCore::write_checkpoint(long step) {
  const long nf = couplb_state_len(ctx);
  std::vector<double> f(nf);
  check(couplb_get_state_f(ctx, f.data()), "state export");
  const std::string fname = checkpoint_prefix + ".0.clbk";
  FILE* fp = std::fopen(fname.c_str(), "wb");
  if (!fp) { say("CoupLB IO: rank 0 cannot open %s for writing\n", fname.c_str()); return; }
  const uint32_t magic = CLBK_MAGIC;
  const int32_t ver = CLBK_VERSION, D = is3d ? 3 : 2, Q = is3d ? 19 : 9;
  const int64_t s64 = step;
  const int32_t dims[3] = {Nx, Ny, is3d ? Nz : 1}, offs[3] = {0, 0, 0}, glob[3] = {Nx, Ny, is3d ? Nz : 1};
  bool ok = std::fwrite(&magic, 4, 1, fp) == 1 && std::fwrite(&ver, 4, 1, fp) == 1 &&
            std::fwrite(&s64, 8, 1, fp) == 1 && std::fwrite(&D, 4, 1, fp) == 1 &&
            std::fwrite(&Q, 4, 1, fp) == 1 && std::fwrite(dims, 4, 3, fp) == 3 &&
            std::fwrite(offs, 4, 3, fp) == 3 && std::fwrite(glob, 4, 3, fp) == 3 &&
            std::fwrite(f.data(), sizeof(double), nf, fp) == (size_t)nf;
  std::fclose(fp);
  if (!ok) say("CoupLB IO: rank 0 write error in %s\n", fname.c_str());
  else say("CoupLB: checkpoint written at step %ld -> %s.*.clbk\n", step, checkpoint_prefix.c_str());
}

// IO::read_checkpoint, couplb_io.h:721-828
void FixCoupLBCore::read_checkpoint() {
  const std::string fname = restart_prefix + ".0.clbk";
  FILE* fp = std::fopen(fname.c_str(), "rb");
  auto bad = [&](const char* why) {
    if (fp) std::fclose(fp);
    say("CoupLB IO: rank 0 %s in %s\n", why, fname.c_str());
    throw fix_error("fix couplb: failed to read checkpoint");
  };
  if (!fp) bad("cannot open");
  uint32_t magic;
  int32_t ver, D, Q, dims[3], offs[3], glob[3];
  int64_t s64;
  if (std::fread(&magic, 4, 1, fp) != 1 || magic != CLBK_MAGIC) bad("bad magic");
  if (std::fread(&ver, 4, 1, fp) != 1 || ver != CLBK_VERSION) bad("unsupported version");
  if (std::fread(&s64, 8, 1, fp) != 1) bad("truncated header (step)");
  if (std::fread(&D, 4, 1, fp) != 1 || std::fread(&Q, 4, 1, fp) != 1) bad("truncated header (D/Q)");
  if (D != (is3d ? 3 : 2) || Q != (is3d ? 19 : 9)) bad("D/Q mismatch");
  if (std::fread(dims, 4, 3, fp) != 3 || std::fread(offs, 4, 3, fp) != 3 || std::fread(glob, 4, 3, fp) != 3)
    bad("truncated header (dims)");
  const int nzl = is3d ? Nz : 1;
  if (dims[0] != Nx || dims[1] != Ny || dims[2] != nzl) bad("local grid mismatch");
  if (offs[0] != 0 || offs[1] != 0 || offs[2] != 0) bad("offset mismatch");
  if (glob[0] != Nx || glob[1] != Ny || glob[2] != nzl) bad("global grid mismatch");
  const long nf = couplb_state_len(ctx);
  std::vector<double> f(nf);
  if ((long)std::fread(f.data(), sizeof(double), nf, fp) != nf) bad("short read");
  std::fclose(fp);
  fp = nullptr;
  check(couplb_set_state_f(ctx, f.data()), "state import");
  say("CoupLB: checkpoint loaded from step %ld <- %s.*.clbk\n", (long)s64, restart_prefix.c_str());
}

// FixCoupLB::end_of_step, fix_couplb.cpp:846-901
int FixCoupLBCore::end_of_step(long step) {
  bool chk = false, out = false, vtk = false, ckpt = false;
  if (check_every > 0 && step >= next_check) { chk = true; next_check = step + check_every; }
  if (output_every > 0 && step >= next_output) { out = true; chk = true; next_output += output_every; }
  if (vtk_every > 0 && step >= next_vtk) { vtk = true; next_vtk = step + vtk_every; }
  if (checkpoint_every > 0 && step >= next_checkpoint) { ckpt = true; next_checkpoint = step + checkpoint_every; }
  if (!chk && !out && !vtk && !ckpt) return 0;
  check(flush_steps(), "LB steps");
  if (chk || out || vtk) check(couplb_refresh_macro(ctx), "refresh");
  if (chk) check_stability_precomputed(step);
  if (out) write_profile(step);
  if (vtk) write_vtk(step);
  if (ckpt) write_checkpoint(step);
  return 0;
}

// The vtk block of end_of_step / setup, fix_couplb.cpp:862-876, 588-603:
// fluid fields, the group's atoms, and both collections.
void FixCoupLBCore::write_vtk(long step) {
  const size_t N = (size_t)Nx * Ny * (is3d ? Nz : 1);
  std::vector<double> r(N), ux(N), uy(N), uz(N), F(3 * N);
  check(couplb_get_macro(ctx, r.data(), ux.data(), uy.data(), uz.data()), "macro fields");
  // Grid::fx holds the reactions spread since the LB step; their periodic
  // ghost images reach the interior at the interval's last sub-step
  // (exchange_forces, fix_couplb.cpp:795-797; setup, 558-563)
  check(couplb_get_node_forces(ctx, F.data(), md_sub_count == 0 ? 1 : 0), "node forces");
  VtiFields vf;
  vf.nx = Nx; vf.ny = Ny; vf.nz = is3d ? Nz : 1;
  vf.rho = r.data(); vf.ux = ux.data(); vf.uy = uy.data(); vf.uz = uz.data(); vf.force = F.data();
  std::string fname, msg;
  if (write_vti(vtk_prefix, step, vf, has_vtk_region ? vtk_region : nullptr, dom.boxlo, dx, vel_scale,
                force_scale, rho, &fname, &msg))
    say("CoupLB: VTK written at step %ld -> %s%s\n", step, fname.c_str(), msg.c_str());
  else
    say("%s\n", msg.c_str());
  if (vtk_solid_on) write_solid_vtk(step);
  vtk_steps.push_back(step);
  if (!write_pvd(vtk_pvd_file, vtk_prefix, vtk_steps, dom.dt, "vti"))
    say("CoupLB IO: cannot open %s\n", vtk_pvd_file.c_str());
  if (vtk_solid_on && !write_pvd(vtk_solid_pvd_file, vtk_prefix + "_solid", vtk_steps, dom.dt, "vtp"))
    say("CoupLB IO: cannot open %s\n", vtk_solid_pvd_file.c_str());
}

// write_solid_vtk, couplb_io.h:304-609: the group's atoms (inside the
// snapped vtk_region) as points, with the requested attributes; an x/y/z
// style triple is one 3-component array when all three are requested.
void FixCoupLBCore::write_solid_vtk(long step) {
  const couplb_atoms& a = atoms_now;
  if (!have_atoms) throw fix_error("fix couplb vtk_solid: no atom arrays (setup / post_force not called)");
  const couplb_atom_extras& e = extras;
  const bool ex = have_extras;
  auto tagval = [&](const void* p, int i) -> double {
    return e.tag_bytes == 8 ? (double)static_cast<const int64_t*>(p)[i] : (double)static_cast<const int*>(p)[i];
  };
  struct Attr {
    const char* name;
    const char* group;  // "" for a scalar
    int comp;
    std::function<double(int)> get;
  };
  std::vector<Attr> reg;
  if (ex && e.tag) reg.push_back({"id", "", 0, [&](int i) { return tagval(e.tag, i); }});
  if (a.type) reg.push_back({"type", "", 0, [&](int i) { return (double)a.type[i]; }});
  if (ex && e.molecule) reg.push_back({"mol", "", 0, [&](int i) { return tagval(e.molecule, i); }});
  if (a.rmass) reg.push_back({"mass", "", 0, [&](int i) { return a.rmass[i]; }});
  else if (a.mass && a.type) reg.push_back({"mass", "", 0, [&](int i) { return a.mass[a.type[i]]; }});
  if (ex && e.radius) reg.push_back({"diameter", "", 0, [&](int i) { return 2.0 * e.radius[i]; }});
  const char* xyz[3] = {"x", "y", "z"};
  const char* vxyz[3] = {"vx", "vy", "vz"};
  const char* fxyz[3] = {"fx", "fy", "fz"};
  for (int c = 0; c < 3; c++) reg.push_back({xyz[c], "position", c, [&, c](int i) { return a.x[3 * i + c]; }});
  for (int c = 0; c < 3; c++) reg.push_back({vxyz[c], "velocity", c, [&, c](int i) { return a.v[3 * i + c]; }});
  for (int c = 0; c < 3; c++) reg.push_back({fxyz[c], "force", c, [&, c](int i) { return a.f[3 * i + c]; }});
  const char* mu[3] = {"mux", "muy", "muz"};
  const char* om[3] = {"omegax", "omegay", "omegaz"};
  const char* tq[3] = {"tqx", "tqy", "tqz"};
  if (ex && e.mu)
    for (int c = 0; c < 3; c++) reg.push_back({mu[c], "dipole", c, [&, c](int i) { return e.mu[4 * i + c]; }});
  if (ex && e.omega)
    for (int c = 0; c < 3; c++) reg.push_back({om[c], "omega", c, [&, c](int i) { return e.omega[3 * i + c]; }});
  if (ex && e.torque)
    for (int c = 0; c < 3; c++) reg.push_back({tq[c], "torque", c, [&, c](int i) { return e.torque[3 * i + c]; }});
  auto is_one = [](const std::string& s, const char* const* names) {
    return s == names[0] || s == names[1] || s == names[2];
  };
  std::vector<const Attr*> req;
  for (const std::string& s : vtk_solid_attrs) {
    if (s == "mol" && !(ex && e.molecule))
      throw fix_error("fix couplb vtk_solid: attribute 'mol' requires a molecular atom style");
    if (s == "diameter" && !(ex && e.radius))
      throw fix_error("fix couplb vtk_solid: attribute 'diameter' requires atom style sphere (radius)");
    if (is_one(s, mu) && !(ex && e.mu))
      throw fix_error("fix couplb vtk_solid: dipole attributes require dipole atom style (mu)");
    if (is_one(s, om) && !(ex && e.omega))
      throw fix_error("fix couplb vtk_solid: omega attributes require atom style sphere (omega)");
    if (is_one(s, tq) && !(ex && e.torque))
      throw fix_error("fix couplb vtk_solid: torque attributes require atom style sphere (torque)");
    if (s == "mass" && !(a.rmass || a.mass))
      throw fix_error("fix couplb vtk_solid: attribute 'mass' requires per-atom mass (rmass) or type mass table");
    if (s == "id" && !(ex && e.tag)) throw fix_error("fix couplb vtk_solid: attribute 'id' needs the atom IDs");
    if (s == "type" && !a.type) throw fix_error("fix couplb vtk_solid: attribute 'type' needs the atom types");
    const Attr* hit = nullptr;
    for (const Attr& r : reg)
      if (s == r.name) { hit = &r; break; }
    if (!hit) throw fix_error("fix couplb vtk_solid: unknown attribute '" + s + "'");
    req.push_back(hit);
  }
  // output arrays in request order; a group becomes one array at its first
  // member when all three components are requested
  struct Out {
    std::string name;
    int ncomp;
    const Attr* comp[3];
  };
  std::vector<Out> outs;
  std::set<std::string> done;
  for (const Attr* r : req) {
    if (r->group[0]) {
      if (done.count(r->group)) continue;
      const Attr* comp[3] = {nullptr, nullptr, nullptr};
      for (const Attr* q : req)
        if (!std::strcmp(q->group, r->group)) comp[q->comp] = q;
      if (comp[0] && comp[1] && comp[2]) {
        outs.push_back({r->group, 3, {comp[0], comp[1], comp[2]}});
        done.insert(r->group);
        continue;
      }
    }
    outs.push_back({r->name, 1, {r, nullptr, nullptr}});
  }
  std::vector<double> pts;
  std::vector<VtpArray> arrays(outs.size());
  for (size_t o = 0; o < outs.size(); o++) {
    arrays[o].name = outs[o].name;
    arrays[o].ncomp = outs[o].ncomp;
  }
  for (int i = 0; i < a.nlocal; i++) {
    if (!(a.mask[i] & a.groupbit)) continue;
    if (has_vtk_region) {
      bool in = true;
      for (int d = 0; d < 3; d++)
        if (a.x[3 * i + d] < vtk_rlo[d] || a.x[3 * i + d] >= vtk_rhi[d]) in = false;
      if (!in) continue;
    }
    for (int d = 0; d < 3; d++) pts.push_back(a.x[3 * i + d]);
    for (size_t o = 0; o < outs.size(); o++)
      for (int d = 0; d < outs[o].ncomp; d++) arrays[o].v.push_back(outs[o].comp[d]->get(i));
  }
  std::string fname;
  if (!write_vtp(vtk_prefix, step, pts, arrays, &fname)) throw fix_error("CoupLB IO: cannot open " + fname);
  say("CoupLB: solid VTK written at step %ld -> %s (%d atoms, %d fields)\n", step, fname.c_str(),
      (int)(pts.size() / 3), (int)outs.size());
}

double FixCoupLBCore::scalar(const std::string& n) const {
  if (n == "tau") return tau;
  if (n == "nu_lb") return nu_lb;
  if (n == "dx") return dx;
  if (n == "dt_LBM") return dt_LBM;
  if (n == "vel_scale") return vel_scale;
  if (n == "force_scale") return force_scale;
  if (n == "lbm_steps") return (double)lbm_step_count;
  if (n == "md_sub_count") return md_sub_count;
  if (n == "md_per_lb") return md_per_lb;
  if (n == "lbm_ms") return couplb_last_kernel_ms(ctx, 0);
  if (n == "ibm_ms") return couplb_last_kernel_ms(ctx, 1);
  return std::nan("");
}

bool FixCoupLBCore::last_profile(double* r, double* u, double* v) const {
  if (!have_profile) return false;
  for (int j = 0; j < Ny; j++) { r[j] = prof_r[j]; u[j] = prof_u[j]; v[j] = prof_v[j]; }
  return true;
}

}  // namespace couplb_host

// ---------------------------------------------------------------------------
// C ABI (include/couplb_fix.h)
// ---------------------------------------------------------------------------
struct couplb_fix {
  couplb_host::FixCoupLBCore* core = nullptr;
  std::string err;
};

namespace {
template <typename F>
int guard(couplb_fix* fx, F&& fn) {
  try {
    fn();
    return COUPLB_OK;
  } catch (const std::exception& e) {
    if (fx) fx->err = e.what();
    return COUPLB_EINVAL;
  }
}
}  // namespace

extern "C" {

int couplb_fix_create(int narg, const char* const* arg, const couplb_domain* dom, couplb_fix** out) {
  if (!out || !dom) return COUPLB_EINVAL;
  auto* fx = new couplb_fix();
  *out = fx;
  return guard(fx, [&] { fx->core = new couplb_host::FixCoupLBCore(narg, arg, *dom); });
}

int couplb_fix_destroy(couplb_fix* fx) {
  if (!fx) return COUPLB_OK;
  delete fx->core;
  delete fx;
  return COUPLB_OK;
}

const char* couplb_fix_last_error(const couplb_fix* fx) { return fx ? fx->err.c_str() : "null fix"; }

int couplb_fix_set_variable_callback(couplb_fix* fx, couplb_var_eval fn, void* user) {
  if (!fx || !fx->core) return COUPLB_EINVAL;
  fx->core->set_variable_callback(fn, user);
  return COUPLB_OK;
}

int couplb_fix_set_domain(couplb_fix* fx, const couplb_domain* dom) {
  if (!fx || !fx->core || !dom) return COUPLB_EINVAL;
  fx->core->set_domain(*dom);
  return COUPLB_OK;
}

int couplb_fix_init(couplb_fix* fx) {
  if (!fx || !fx->core) return COUPLB_EINVAL;
  return guard(fx, [&] { fx->core->init(); });
}
int couplb_fix_setup(couplb_fix* fx, const couplb_atoms* a) {
  if (!fx || !fx->core || !a) return COUPLB_EINVAL;
  return guard(fx, [&] { fx->core->setup(*a); });
}
int couplb_fix_post_force(couplb_fix* fx, const couplb_atoms* a) {
  if (!fx || !fx->core || !a) return COUPLB_EINVAL;
  return guard(fx, [&] { fx->core->post_force(*a); });
}
int couplb_fix_end_of_step(couplb_fix* fx, long step) {
  if (!fx || !fx->core) return COUPLB_EINVAL;
  return guard(fx, [&] { fx->core->end_of_step(step); });
}
int couplb_fix_set_atom_extras(couplb_fix* fx, const couplb_atom_extras* e) {
  if (!fx || !fx->core || !e) return COUPLB_EINVAL;
  if (e->tag_bytes != 4 && e->tag_bytes != 8 && (e->tag || e->molecule)) {
    fx->err = "couplb_atom_extras: tag_bytes must be 4 or 8";
    return COUPLB_EINVAL;
  }
  fx->core->set_atom_extras(*e);
  return COUPLB_OK;
}
int couplb_fix_setmask(void) { return couplb_host::FixCoupLBCore::setmask(); }

double couplb_fix_scalar(const couplb_fix* fx, const char* name) {
  return (fx && fx->core) ? fx->core->scalar(name) : std::nan("");
}
int couplb_fix_last_diag(const couplb_fix* fx, double out[6]) {
  if (!fx || !fx->core) return COUPLB_EINVAL;
  for (int i = 0; i < 6; i++) out[i] = fx->core->last_diag()[i];
  return COUPLB_OK;
}
int couplb_fix_last_profile(const couplb_fix* fx, double* r, double* u, double* v) {
  if (!fx || !fx->core) return COUPLB_EINVAL;
  return fx->core->last_profile(r, u, v) ? COUPLB_OK : COUPLB_ESTATE;
}
void* couplb_fix_context(couplb_fix* fx) { return (fx && fx->core) ? fx->core->context() : nullptr; }
const char* couplb_fix_log(const couplb_fix* fx) { return (fx && fx->core) ? fx->core->log().c_str() : ""; }

}  // extern "C"
