// vtk_io.hpp -- VTK output of `fix couplb`: the fluid fields as ImageData
// (.vti), the group's atoms as PolyData (.vtp) and ParaView collections
// (.pvd).  The files have the layout CoupLB writes (couplb_io.h: write_vtk
// 42-275, write_solid_vtk 304-609, write_pvd_impl 614-647): ASCII XML,
// values printed with %.8e, cell data rho / velocity_phys / type /
// force_phys in that order, x fastest.
#pragma once
#include <string>
#include <vector>

namespace couplb_host {

// Fluid fields of an nx*ny*nz interior grid (x fastest), lattice units.
struct VtiFields {
  int nx = 0, ny = 0, nz = 1;
  const double* rho = nullptr;
  const double* ux = nullptr;
  const double* uy = nullptr;
  const double* uz = nullptr;
  const double* force = nullptr;  // 3 per node (fx, fy, fz); null reads as zero
};

// One .vti of the node box [region[0], region[1]) x [region[2], region[3])
// x [region[4], region[5]) (null region: the whole grid), named
// prefix_STEP.vti.  Origin = domain_lo + region lower corner * dx, spacing
// dx.  Fields are scaled to physical units: rho * rho_scale,
// u * vel_scale, F * force_scale; type is 0 (interior nodes are fluid).
// Returns false with *msg set when the region is empty or the file cannot
// be opened (the reference prints and carries on, couplb_io.h:61-71, 110).
bool write_vti(const std::string& prefix, long step, const VtiFields& f, const int* region,
               const double domain_lo[3], double dx, double vel_scale, double force_scale,
               double rho_scale, std::string* fname, std::string* msg);

// Collection file listing prefix_STEP.ext for every step, timestep = step * dt,
// file names without their directory (couplb_io.h:614-647).
bool write_pvd(const std::string& pvd_file, const std::string& prefix, const std::vector<long>& steps,
               double dt, const char* ext);

// One named PointData array: ncomp (1 or 3) values per point.
struct VtpArray {
  std::string name;
  int ncomp = 1;
  std::vector<double> v;
};

// prefix_solid_STEP.vtp with the points xyz (3 per point), one vertex per
// point and the given arrays.  Returns false when the file cannot be opened.
bool write_vtp(const std::string& prefix, long step, const std::vector<double>& xyz,
               const std::vector<VtpArray>& arrays, std::string* fname);

}  // namespace couplb_host
