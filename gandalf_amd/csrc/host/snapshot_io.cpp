// Column-format snapshots of the device-resident state.
//
// Simulation::WriteColumnSnapshotFile / ReadColumnSnapshotFile
// (src/Common/SimulationIO.hpp:444-548, :193-272), 3D, dimensionless units (every outscale is
// 1): a header of Nhydro, Nstar, ndim and t, then one line per hydro particle with r, v, m, h,
// rho, u and one per star with r, v, m, h, 0, 0, written with the default formatting of a C++
// stream (6 significant digits) and "   " between values, as the reference writes them.
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "../gdf_internal.h"

namespace {

void put_line(std::ofstream& f, const double* r, const double* v, double m, double h, double rho,
              double u, bool star) {
  f << r[0] << "   " << r[1] << "   " << r[2] << "   " << v[0] << "   " << v[1] << "   "
    << v[2] << "   " << m << "   " << h << "   ";
  if (star) f << 0.0 << "   " << 0.0 << std::endl;
  else f << rho << "   " << u << std::endl;
}

}  // namespace

extern "C" {

int gdf_format_column(const char* path, int nhydro, int nstar, double t, const double* r,
                      const double* v, const double* m, const double* h, const double* rho,
                      const double* u) {
  if (!path || nhydro < 0 || nstar < 0 || (nhydro + nstar > 0 && (!r || !v || !m || !h)))
    return GDF_ERR_INVALID;
  if (nhydro > 0 && (!rho || !u)) return GDF_ERR_INVALID;
  std::ofstream f(path);
  if (!f) return GDF_ERR_INVALID;
  f << nhydro << std::endl << nstar << std::endl << 3 << std::endl << t << std::endl;
  for (int i = 0; i < nhydro + nstar; i++) {
    const bool star = i >= nhydro;
    put_line(f, r + 3*(size_t) i, v + 3*(size_t) i, m[i], h[i], star ? 0.0 : rho[i],
             star ? 0.0 : u[i], star);
  }
  return f.good() ? GDF_OK : GDF_ERR_INVALID;
}

int gdf_write_snapshot(gdf_ctx* c, const char* path, const char* format) {
  if (!c || !path || !format || c->N <= 0) return GDF_ERR_INVALID;
  if (std::string(format) != "column")
    return gdf::fail(c, GDF_ERR_INVALID, std::string("unsupported snapshot format ") + format);
  const size_t n = (size_t) c->N;
  std::vector<double> r(3*n), v(3*n), m(n), h(n), rho(n), u(n);
  const char* names[] = {"r", "v", "m", "h", "rho", "u"};
  double* bufs[] = {r.data(), v.data(), m.data(), h.data(), rho.data(), u.data()};
  for (int k = 0; k < 6; k++) {
    const int rc = gdf_get_field(c, names[k], bufs[k]);
    if (rc != GDF_OK) return rc;
  }
  // an N-body context holds stars only (NbodySimulation writes them as the star block)
  const int nh = c->phys.nbody ? 0 : c->N, ns = c->phys.nbody ? c->N : 0;
  const int rc = gdf_format_column(path, nh, ns, c->t, r.data(), v.data(), m.data(), h.data(),
                                   rho.data(), u.data());
  return rc == GDF_OK ? rc : gdf::fail(c, rc, std::string("cannot write ") + path);
}

int gdf_read_snapshot(gdf_ctx* c, const char* path, const char* format) {
  if (!c || !path || !format) return GDF_ERR_INVALID;
  if (std::string(format) != "column")
    return gdf::fail(c, GDF_ERR_INVALID, std::string("unsupported snapshot format ") + format);
  std::ifstream f(path);
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, std::string("cannot read ") + path);
  int nh = 0, ns = 0, ndim = 0;
  double t = 0.0;
  f >> nh >> ns >> ndim >> t;
  if (!f || ndim != 3 || nh < 0 || ns < 0 || nh + ns <= 0)
    return gdf::fail(c, GDF_ERR_INVALID, "bad column snapshot header");
  if (c->phys.nbody ? nh != 0 : ns != 0)
    return gdf::fail(c, GDF_ERR_INVALID, c->phys.nbody ? "N-body context: hydro particles in snapshot"
                                                       : "hydro context: star particles in snapshot");
  const size_t n = (size_t) (nh + ns);
  std::vector<double> r(3*n), v(3*n), m(n), h(n), rho(n), u(n);
  for (size_t i = 0; i < n; i++) {
    f >> r[3*i] >> r[3*i + 1] >> r[3*i + 2] >> v[3*i] >> v[3*i + 1] >> v[3*i + 2] >> m[i] >> h[i]
      >> rho[i] >> u[i];
    if (!f) return gdf::fail(c, GDF_ERR_INVALID, "truncated column snapshot");
  }
  int rc = gdf_set_particles(c, (int) n, r.data(), v.data(), m.data(), h.data(), u.data());
  if (rc != GDF_OK) return rc;
  if (!c->phys.nbody) rc = gdf_set_field(c, "rho", rho.data());
  c->t = t;
  return rc;
}

}  // extern "C"
