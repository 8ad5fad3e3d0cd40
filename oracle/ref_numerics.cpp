// ref_numerics.cpp -- exposes the reference's own numerics headers through
// the oracle's clo_ops table.  TEST INFRASTRUCTURE ONLY.
//
// Built by oracle/Makefile directly against the unmodified headers under
// /root/reference/CoupLB (couplb_lattice.h, couplb_collision.h,
// couplb_streaming.h, couplb_boundary.h, couplb_ibm.h) into
// oracle/_ref/libcouplb_ref.so.  Nothing here restates the numerics: every
// primitive forwards to Grid<L>, BGK<L>, Streaming<L>, Boundary<L> or IBM<L>.
// The orchestration that fix_couplb.cpp performs (it needs LAMMPS headers and
// does not build here) is the oracle's clo_fix layer, which drives this table.
//
// Streaming is configured single-rank and self-periodic in every dimension,
// so only the copy/fold branches run (couplb_streaming.h:176-182, 196-202,
// 217-225) and no MPI routine is entered; libmpi from the image is linked
// only because the instantiated templates reference MPI_Sendrecv.
#include "couplb_lattice.h"
#include "couplb_collision.h"
#include "couplb_streaming.h"
#include "couplb_boundary.h"
#include "couplb_ibm.h"

#include "couplb_oracle.h"

#include <cstring>

using namespace LAMMPS_NS::CoupLB;

namespace {

struct RefBase {
  virtual ~RefBase() {}
  int D;
};

template <typename L>
struct Ref : RefBase {
  Grid<L> g;
  BGK<L> bgk;
  Streaming<L> st;
};

template <typename L>
Ref<L>* as(void* p) { return static_cast<Ref<L>*>(static_cast<RefBase*>(p)); }

template <typename L>
void* create_t(const clo_grid_params* p) {
  auto* r = new Ref<L>();
  r->D = L::D;
  const int nz = (L::D == 3) ? p->nz : 1;
  r->g.allocate(p->nx, p->ny, nz, 1.0, 0, 0, 0, p->nx, p->ny, nz);
  r->g.init_equilibrium(p->rho_lb, 0, 0, 0);
  r->bgk.set_tau(p->tau);
  r->st.set_periodic(true, true, L::D == 3);
  r->st.set_nprocs(1, 1, 1);
  r->st.allocate_buffers(r->g, p->rho_lb);
  // FixCoupLB::setup_boundaries, fix_couplb.cpp:522-539 (single rank)
  const int* w0 = p->wall[0];
  const int* w1 = p->wall[1];
  const int* w2 = p->wall[2];
  if (w0[0] || w0[1]) Boundary<L>::set_walls_x(r->g, w0[0] > 0, w0[1] > 0, w0[0], w0[1]);
  if (w1[0] || w1[1]) Boundary<L>::set_walls_y(r->g, w1[0] > 0, w1[1] > 0, w1[0], w1[1]);
  if (L::D == 3 && (w2[0] || w2[1]))
    Boundary<L>::set_walls_z(r->g, w2[0] > 0, w2[1] > 0, w2[0], w2[1]);
  // The reference has one velocity for all type-2 nodes; use the first
  // moving face's velocity (all faces carry the same one in its configs).
  double v[3] = {0, 0, 0};
  bool found = false;
  for (int d = 0; d < L::D && !found; d++)
    for (int s = 0; s < 2 && !found; s++)
      if (p->wall[d][s] == 2) {
        for (int c = 0; c < 3; c++) v[c] = p->face_vel[d][s][c];
        found = true;
      }
  Boundary<L>::set_wall_velocity(r->g, 2, v[0], v[1], (L::D == 3) ? v[2] : 0.0);
  r->st.precompute_wall_flags(r->g);
  return static_cast<RefBase*>(r);
}

void* r_create(const clo_grid_params* p) {
  return p->D == 3 ? create_t<D3Q19>(p) : create_t<D2Q9>(p);
}
void r_destroy(void* p) { delete static_cast<RefBase*>(p); }

#define DISPATCH(p, CALL)                                              \
  do {                                                                 \
    if (static_cast<RefBase*>(p)->D == 3) { auto* r = as<D3Q19>(p); CALL; } \
    else { auto* r = as<D2Q9>(p); CALL; }                             \
  } while (0)

template <typename L>
double* dfield_t(Ref<L>* r, int w) {
  auto& g = r->g;
  switch (w) {
    case CLO_F_RHO: return g.rho.data();   case CLO_F_UX: return g.ux.data();
    case CLO_F_UY: return g.uy.data();     case CLO_F_UZ: return g.uz.data();
    case CLO_F_FX: return g.fx.data();     case CLO_F_FY: return g.fy.data();
    case CLO_F_FZ: return g.fz.data();     case CLO_F_BCUX: return g.bc_ux.data();
    case CLO_F_BCUY: return g.bc_uy.data(); case CLO_F_BCUZ: return g.bc_uz.data();
    case CLO_F_F: return g.f.data();       case CLO_F_FBUF: return g.f_buf.data();
  }
  return nullptr;
}

double* r_dfield(void* p, int w) { DISPATCH(p, return dfield_t(r, w)); return nullptr; }
int* r_ifield(void* p, int w) {
  DISPATCH(p, return (w == CLO_I_TYPE) ? r->g.type.data() : r->g.wall_dim.data());
  return nullptr;
}
void r_dims(void* p, int gd[3], long* nt, int* Q) {
  DISPATCH(p, { gd[0] = r->g.gx; gd[1] = r->g.gy; gd[2] = r->g.gz; *nt = r->g.ntotal; *Q = r->g.Q; });
}
int r_macro(void* p, int guo) {
  DISPATCH(p, { r->g.compute_macroscopic(guo != 0); return r->g.rho_clamp_count; });
  return 0;
}
void r_collide(void* p) { DISPATCH(p, r->bgk.collide(r->g)); }
void r_exchange(void* p) { DISPATCH(p, r->st.exchange(r->g)); }
void r_stream(void* p) { DISPATCH(p, r->st.stream(r->g)); }
void r_exchange_velocity(void* p) { DISPATCH(p, r->st.exchange_velocity(r->g)); }
void r_exchange_forces(void* p) { DISPATCH(p, r->st.exchange_forces(r->g)); }
void r_clear_forces(void* p) { DISPATCH(p, r->g.clear_forces()); }

DeltaKernel kern_of(int k) { return k == CLO_ROMA ? DeltaKernel::ROMA : DeltaKernel::PESKIN4; }


template <typename L>
void interp_t(Ref<L>* r, double x, double y, double z, const double lo[3],
              double inv_dx, int kern, double out[3]) {
  auto u = IBM<L>::interpolate(r->g, x, y, z, lo, inv_dx, kern_of(kern));
  out[0] = u[0]; out[1] = u[1]; out[2] = u[2];
}
void r_interp(void* p, double x, double y, double z, const double lo[3],
              double inv_dx, int kern, double out[3]) {
  DISPATCH(p, interp_t(r, x, y, z, lo, inv_dx, kern, out));
}
template <typename L>
void spread_t(Ref<L>* r, double x, double y, double z, double fx, double fy,
              double fz, double dvl, const double lo[3], double inv_dx,
              double dxp, int kern) {
  IBM<L>::spread(r->g, x, y, z, fx, fy, fz, dvl, lo, inv_dx, dxp, kern_of(kern));
}
void r_spread(void* p, double x, double y, double z, double fx, double fy,
              double fz, double dvl, const double lo[3], double inv_dx,
              double dxp, int kern) {
  DISPATCH(p, spread_t(r, x, y, z, fx, fy, fz, dvl, lo, inv_dx, dxp, kern));
}
double r_max_velocity(void* p) { DISPATCH(p, return r->g.max_velocity()); return 0; }
void r_diagnostics(void* p, double out[4]) {
  DISPATCH(p, r->g.compute_diagnostics(out[0], out[1], out[2], out[3]));
}
double r_delta(int kern, double x) {
  return kern == CLO_ROMA ? delta_roma(x) : delta_peskin4(x);
}

const clo_ops REF_OPS = {
  "reference-headers",
  r_create, r_destroy, r_dfield, r_ifield, r_dims,
  r_macro, r_collide, r_exchange, r_stream, r_exchange_velocity,
  r_exchange_forces, r_clear_forces, r_interp, r_spread,
  r_max_velocity, r_diagnostics, r_delta
};
}  // namespace

extern "C" const clo_ops* clr_ops(void) { return &REF_OPS; }
