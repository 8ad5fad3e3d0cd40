// vtk_io.cpp -- VTK writers of `fix couplb` (see vtk_io.hpp).  One device
// grid holds the whole domain, so there is nothing to gather: the caller
// hands over the downloaded fields and the writer streams them out.
#include "vtk_io.hpp"

#include <cstdio>
#include <cstring>

namespace couplb_host {

namespace {

std::string step_name(const std::string& prefix, long step, const char* ext) {
  char buf[64];
  std::snprintf(buf, sizeof(buf), "_%06ld.%s", step, ext);
  return prefix + buf;
}

}  // namespace

bool write_vti(const std::string& prefix, long step, const VtiFields& f, const int* region,
               const double domain_lo[3], double dx, double vel_scale, double force_scale,
               double rho_scale, std::string* fname, std::string* msg) {
  const int ilo = region ? region[0] : 0, ihi = region ? region[1] : f.nx;
  const int jlo = region ? region[2] : 0, jhi = region ? region[3] : f.ny;
  const int klo = region ? region[4] : 0, khi = region ? region[5] : f.nz;
  const int Rx = ihi - ilo, Ry = jhi - jlo, Rz = khi - klo;
  if (Rx <= 0 || Ry <= 0 || Rz <= 0) {
    char b[256];
    std::snprintf(b, sizeof(b),
                  "CoupLB::IO::write_vtk: invalid region [%d,%d) x [%d,%d) x [%d,%d); extents must be positive",
                  ilo, ihi, jlo, jhi, klo, khi);
    if (msg) *msg = b;
    return false;
  }
  const std::string name = step_name(prefix, step, "vti");
  if (fname) *fname = name;
  FILE* fp = std::fopen(name.c_str(), "w");
  if (!fp) {
    if (msg) *msg = "CoupLB IO: cannot open " + name;
    return false;
  }
  const double o[3] = {domain_lo[0] + ilo * dx, domain_lo[1] + jlo * dx, domain_lo[2] + klo * dx};
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"ImageData\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fp,
               "  <ImageData WholeExtent=\"0 %d 0 %d 0 %d\" Origin=\"%.8e %.8e %.8e\" "
               "Spacing=\"%.8e %.8e %.8e\">\n",
               Rx, Ry, Rz, o[0], o[1], o[2], dx, dx, dx);
  std::fprintf(fp, "    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", Rx, Ry, Rz);
  std::fprintf(fp, "      <CellData Scalars=\"rho\" Vectors=\"velocity_phys\">\n");
  // cells of the region in file order (x fastest), as grid node indices
  auto each = [&](auto&& fn) {
    for (int k = klo; k < khi; k++)
      for (int j = jlo; j < jhi; j++)
        for (int i = ilo; i < ihi; i++)
          fn((size_t)i + (size_t)f.nx * ((size_t)j + (size_t)f.ny * (size_t)k));
  };
  std::fprintf(fp, "        <DataArray type=\"Float64\" Name=\"rho\" format=\"ascii\">\n");
  each([&](size_t n) { std::fprintf(fp, "%.8e\n", f.rho[n] * rho_scale); });
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Float64\" Name=\"velocity_phys\" NumberOfComponents=\"3\" "
                   "format=\"ascii\">\n");
  each([&](size_t n) {
    std::fprintf(fp, "%.8e %.8e %.8e\n", f.ux[n] * vel_scale, f.uy[n] * vel_scale, f.uz[n] * vel_scale);
  });
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"type\" format=\"ascii\">\n");
  each([&](size_t) { std::fprintf(fp, "%d\n", 0); });
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Float64\" Name=\"force_phys\" NumberOfComponents=\"3\" "
                   "format=\"ascii\">\n");
  each([&](size_t n) {
    const double* F = f.force ? f.force + 3 * n : nullptr;
    std::fprintf(fp, "%.8e %.8e %.8e\n", F ? F[0] * force_scale : 0.0, F ? F[1] * force_scale : 0.0,
                 F ? F[2] * force_scale : 0.0);
  });
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </CellData>\n");
  std::fprintf(fp, "    </Piece>\n");
  std::fprintf(fp, "  </ImageData>\n");
  std::fprintf(fp, "</VTKFile>\n");
  std::fclose(fp);
  if (msg) {
    char b[128];
    std::snprintf(b, sizeof(b), " (%dx%dx%d region)", Rx, Ry, Rz);
    *msg = b;
  }
  return true;
}

bool write_pvd(const std::string& pvd_file, const std::string& prefix, const std::vector<long>& steps,
               double dt, const char* ext) {
  FILE* fp = std::fopen(pvd_file.c_str(), "w");
  if (!fp) return false;
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fp, "  <Collection>\n");
  for (long s : steps) {
    const std::string name = step_name(prefix, s, ext);
    const size_t cut = name.find_last_of('/');
    const std::string base = cut == std::string::npos ? name : name.substr(cut + 1);
    std::fprintf(fp, "    <DataSet timestep=\"%.8e\" file=\"%s\"/>\n", (double)s * dt, base.c_str());
  }
  std::fprintf(fp, "  </Collection>\n");
  std::fprintf(fp, "</VTKFile>\n");
  std::fclose(fp);
  return true;
}

bool write_vtp(const std::string& prefix, long step, const std::vector<double>& xyz,
               const std::vector<VtpArray>& arrays, std::string* fname) {
  const std::string name = step_name(prefix + "_solid", step, "vtp");
  if (fname) *fname = name;
  FILE* fp = std::fopen(name.c_str(), "w");
  if (!fp) return false;
  const int np = (int)(xyz.size() / 3);
  std::fprintf(fp, "<?xml version=\"1.0\"?>\n");
  std::fprintf(fp, "<VTKFile type=\"PolyData\" version=\"1.0\" byte_order=\"LittleEndian\">\n");
  std::fprintf(fp, "  <PolyData>\n");
  std::fprintf(fp,
               "    <Piece NumberOfPoints=\"%d\" NumberOfVerts=\"%d\" NumberOfLines=\"0\" "
               "NumberOfStrips=\"0\" NumberOfPolys=\"0\">\n",
               np, np);
  std::fprintf(fp, "      <Points>\n");
  std::fprintf(fp, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int i = 0; i < np; i++) std::fprintf(fp, "%.8e %.8e %.8e\n", xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Points>\n");
  std::fprintf(fp, "      <Verts>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n");
  for (int i = 0; i < np; i++) std::fprintf(fp, "%d\n", i);
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n");
  for (int i = 0; i < np; i++) std::fprintf(fp, "%d\n", i + 1);
  std::fprintf(fp, "        </DataArray>\n");
  std::fprintf(fp, "      </Verts>\n");
  std::fprintf(fp, "      <PointData>\n");
  for (const VtpArray& a : arrays) {
    if (a.ncomp == 1) {
      std::fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", a.name.c_str());
      for (int i = 0; i < np; i++) std::fprintf(fp, "%.8e\n", a.v[i]);
    } else {
      std::fprintf(fp, "        <DataArray type=\"Float64\" Name=\"%s\" NumberOfComponents=\"3\" format=\"ascii\">\n",
                   a.name.c_str());
      for (int i = 0; i < np; i++) std::fprintf(fp, "%.8e %.8e %.8e\n", a.v[3 * i], a.v[3 * i + 1], a.v[3 * i + 2]);
    }
    std::fprintf(fp, "        </DataArray>\n");
  }
  std::fprintf(fp, "      </PointData>\n");
  std::fprintf(fp, "    </Piece>\n");
  std::fprintf(fp, "  </PolyData>\n");
  std::fprintf(fp, "</VTKFile>\n");
  std::fclose(fp);
  return true;
}

}  // namespace couplb_host
