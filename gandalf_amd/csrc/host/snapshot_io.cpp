// Snapshots of the device-resident state: the reference's column text format and its Seren
// unformatted binary format (the default out_file_form, Parameters.cpp:175-176).
//
// Simulation::WriteColumnSnapshotFile / ReadColumnSnapshotFile
// (src/Common/SimulationIO.hpp:444-548, :193-272), 3D, dimensionless units (every outscale is
// 1): a header of Nhydro, Nstar, ndim and t, then one line per hydro particle with r, v, m, h,
// rho, u and one per star with r, v, m, h, 0, 0, written with the default formatting of a C++
// stream (6 significant digits) and "   " between values, as the reference writes them.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <new>
#include <iomanip>
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

// Seren unformatted layout (SimulationIO.hpp:2009-2249; BinaryIO.h BinaryWriter writes each
// value's bytes as they lie in memory): a 20-character space-padded tag, the FLOAT size and
// three dimension counts, 50 int, 50 long, 50 FLOAT and 50 double header words, the unit and
// array names as 20-character fields, five ints per array (width, 1, count, type code, unit
// code), then the arrays; stars are one "sink_v1" block of 12 + 2*ndim FLOATs each.
constexpr int kSuStr = 20;
constexpr int kSinkLen = 12 + 2*3;
const char* const kSuTag = "SERENBINARYDUMPV3";

void su_str(std::ofstream& f, const char* s) {
  char b[kSuStr];
  std::memset(b, ' ', kSuStr);
  std::memcpy(b, s, std::min<size_t>(std::strlen(s), kSuStr));
  f.write(b, kSuStr);
}

template <class T> void su_put(std::ofstream& f, T v) { f.write((const char*) &v, sizeof(T)); }
template <class T> bool su_get(std::ifstream& f, T& v) { return (bool) f.read((char*) &v, sizeof(T)); }

std::string su_name(std::ifstream& f) {
  char b[kSuStr];
  if (!f.read(b, kSuStr)) return std::string();
  std::string s(b, kSuStr);
  const size_t a = s.find_first_not_of(" \t"), e = s.find_last_not_of(" \t");
  return a == std::string::npos ? std::string() : s.substr(a, e - a + 1);
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

int gdf_format_su(const char* path, int nhydro, int nstar, const int64_t* counts3,
                  const double* times4, double h_fac, const int* porig, const double* r,
                  const double* v, const double* m, const double* h, const double* rho,
                  const double* u) {
  if (!path || !counts3 || !times4 || nhydro < 0 || nstar < 0 ||
      (nhydro + nstar > 0 && (!r || !v || !m || !h)) || (nhydro > 0 && (!rho || !u)))
    return GDF_ERR_INVALID;
  std::ofstream f(path, std::ios::binary);
  if (!f) return GDF_ERR_INVALID;
  int idata[50] = {};
  long ilpdata[50] = {};
  double rdata[50] = {}, ddata[50] = {};
  // every particle here is gas: the type counts are icm, gas, cdm, dust in idata[3..6]
  idata[0] = nhydro; idata[1] = nstar; idata[4] = nhydro;
  struct Arr { const char* id; int w, type, unit; };
  const Arr hyd[] = {{"porig", 1, 2, 0}, {"r", 3, 4, 1}, {"m", 1, 4, 2}, {"h", 1, 4, 1},
                     {"v", 3, 4, 4}, {"rho", 1, 4, 6}, {"u", 1, 4, 20}};
  const int ndata = (nhydro > 0 ? 7 : 0) + (nstar > 0 ? 1 : 0);
  idata[19] = 0;                       // no unit names: dimensionless
  idata[20] = ndata;
  ilpdata[0] = (long) counts3[0]; ilpdata[1] = (long) counts3[1]; ilpdata[10] = (long) counts3[2];
  rdata[0] = h_fac;
  ddata[0] = times4[0]; ddata[1] = times4[1]; ddata[2] = times4[2]; ddata[10] = times4[3];
  su_str(f, kSuTag);
  su_put(f, (int) sizeof(double));
  for (int k = 0; k < 3; k++) su_put(f, 3);
  for (int k = 0; k < 50; k++) su_put(f, idata[k]);
  for (int k = 0; k < 50; k++) su_put(f, ilpdata[k]);
  for (int k = 0; k < 50; k++) su_put(f, rdata[k]);
  for (int k = 0; k < 50; k++) su_put(f, ddata[k]);
  if (nhydro > 0) for (const Arr& a : hyd) su_str(f, a.id);
  if (nstar > 0) su_str(f, "sink_v1");
  if (nhydro > 0)
    for (const Arr& a : hyd) {
      const int td[5] = {a.w, 1, nhydro, a.type, a.unit};
      for (int k = 0; k < 5; k++) su_put(f, td[k]);
    }
  if (nstar > 0) {
    const int td[5] = {1, 1, nstar, 7, 0};
    for (int k = 0; k < 5; k++) su_put(f, td[k]);
  }
  const size_t n = (size_t) nhydro;
  if (nhydro > 0) {
    for (size_t i = 0; i < n; i++) su_put(f, porig ? porig[i] : (int) i);
    for (size_t i = 0; i < 3*n; i++) su_put(f, r[i]);
    for (size_t i = 0; i < n; i++) su_put(f, m[i]);
    for (size_t i = 0; i < n; i++) su_put(f, h[i]);
    for (size_t i = 0; i < 3*n; i++) su_put(f, v[i]);
    for (size_t i = 0; i < n; i++) su_put(f, rho[i]);
    for (size_t i = 0; i < n; i++) su_put(f, u[i]);
  }
  if (nstar > 0) {
    const int head[6] = {2, 2, 0, kSinkLen, 0, 0};
    for (int k = 0; k < 6; k++) su_put(f, head[k]);
    for (int i = 0; i < nstar; i++) {
      const size_t j = n + (size_t) i;
      su_put(f, 1); su_put(f, 1); su_put(f, i + 1); su_put(f, 0);
      double sd[kSinkLen] = {};
      for (int k = 0; k < 3; k++) { sd[1 + k] = r[3*j + k]; sd[4 + k] = v[3*j + k]; }
      sd[7] = m[j]; sd[8] = h[j];          // sd[9], the star's radius, is 0 here
      for (int k = 0; k < kSinkLen; k++) su_put(f, sd[k]);
    }
  }
  return f.good() ? GDF_OK : GDF_ERR_INVALID;
}

// Seren formatted layout (WriteSerenFormSnapshotFile, SimulationIO.hpp:993-1222, with its
// formatted_output stream, formatted_output.h:27-80): the same header words and arrays as the
// unformatted file, one value per line, integers right-aligned in 2 (the first four lines), 10
// and, in the star records, 8 columns, reals in scientific notation with 10 digits in 18.
int gdf_format_sf(const char* path, int nhydro, int nstar, const int64_t* counts3,
                  const double* times4, double h_fac, const int* porig, const double* r,
                  const double* v, const double* m, const double* h, const double* rho,
                  const double* u) {
  if (!path || !counts3 || !times4 || nhydro < 0 || nstar < 0 ||
      (nhydro + nstar > 0 && (!r || !v || !m || !h)) || (nhydro > 0 && (!rho || !u)))
    return GDF_ERR_INVALID;
  std::ofstream f(path);
  if (!f) return GDF_ERR_INVALID;
  f.setf(std::ios::scientific, std::ios::floatfield);
  auto I = [&f](long long x, int w) { f << std::setw(w) << x; };
  auto D = [&f](double x) { f << std::setw(18) << std::setprecision(10) << x; };
  const int ndata = (nhydro > 0 ? 7 : 0) + (nstar > 0 ? 1 : 0);
  long long idata[50] = {}, ilpdata[50] = {};
  double rdata[50] = {}, ddata[50] = {};
  idata[0] = nhydro; idata[1] = nstar; idata[4] = nhydro; idata[20] = ndata;
  ilpdata[0] = counts3[0]; ilpdata[1] = counts3[1]; ilpdata[10] = counts3[2];
  rdata[0] = h_fac;
  ddata[0] = times4[0]; ddata[1] = times4[1]; ddata[2] = times4[2]; ddata[10] = times4[3];
  f << "SERENASCIIDUMPV2" << std::endl;
  for (int x : {4, 3, 3, 3}) { I(x, 2); f << std::endl; }
  for (long long x : idata) { I(x, 10); f << std::endl; }
  for (long long x : ilpdata) { I(x, 10); f << std::endl; }
  for (double x : rdata) { D(x); f << std::endl; }
  for (double x : ddata) { D(x); f << std::endl; }
  struct Arr { const char* id; int w, type, unit; };
  const Arr hyd[] = {{"porig", 1, 2, 0}, {"r", 3, 4, 1}, {"m", 1, 4, 2}, {"h", 1, 4, 1},
                     {"v", 3, 4, 4}, {"rho", 1, 4, 6}, {"u", 1, 4, 20}};
  if (nhydro > 0) for (const Arr& a : hyd) f << a.id << std::endl;
  if (nstar > 0) f << "sink_v1" << std::endl;
  if (nhydro > 0)
    for (const Arr& a : hyd) {
      for (int x : {a.w, 1, nhydro, a.type, a.unit}) I(x, 10);
      f << std::endl;
    }
  if (nstar > 0) { for (int x : {1, 1, nstar, 7, 0}) I(x, 10); f << std::endl; }
  const size_t n = (size_t) nhydro;
  auto vecs = [&](const double* a) {
    for (size_t i = 0; i < n; i++) { D(a[3*i]); D(a[3*i + 1]); D(a[3*i + 2]); f << std::endl; }
  };
  auto scal = [&](const double* a) { for (size_t i = 0; i < n; i++) { D(a[i]); f << std::endl; } };
  if (nhydro > 0) {
    for (size_t i = 0; i < n; i++) { I(porig ? porig[i] : (long long) i, 10); f << std::endl; }
    vecs(r); scal(m); scal(h); vecs(v); scal(rho); scal(u);
  }
  if (nstar > 0) {
    for (int x : {2, 2, 0, kSinkLen, 0, 0}) I(x, 10);
    f << std::endl;
    for (int i = 0; i < nstar; i++) {
      const size_t j = n + (size_t) i;
      f << "TT" << std::endl;
      I(i + 1, 8); I(0, 8); f << std::endl;
      double sd[kSinkLen] = {};
      for (int k = 0; k < 3; k++) { sd[1 + k] = r[3*j + k]; sd[4 + k] = v[3*j + k]; }
      sd[7] = m[j]; sd[8] = h[j];
      for (double x : sd) D(x);
      f << std::endl;
    }
  }
  return f.good() ? GDF_OK : GDF_ERR_INVALID;
}

// Seren lite layout (WriteSerenLiteSnapshotFile, SimulationIO.hpp:2264-2497): the unformatted
// file's header with a FLOAT size of 4, single-precision header reals, no porig or velocity
// arrays, r, m, h, rho and u as floats, and 4-byte star records; its header keeps the snapshot
// count and step count only (no lite counters).
int gdf_format_slite(const char* path, int nhydro, int nstar, const int64_t* counts2,
                     const double* times3, double h_fac, const double* r, const double* v,
                     const double* m, const double* h, const double* rho, const double* u) {
  if (!path || !counts2 || !times3 || nhydro < 0 || nstar < 0 ||
      (nhydro + nstar > 0 && (!r || !v || !m || !h)) || (nhydro > 0 && (!rho || !u)))
    return GDF_ERR_INVALID;
  std::ofstream f(path, std::ios::binary);
  if (!f) return GDF_ERR_INVALID;
  int idata[50] = {};
  long ilpdata[50] = {};
  float rdata[50] = {};
  double ddata[50] = {};
  idata[0] = nhydro; idata[1] = nstar; idata[4] = nhydro;
  struct Arr { const char* id; int w, unit; };
  const Arr hyd[] = {{"r", 3, 1}, {"m", 1, 2}, {"h", 1, 1}, {"rho", 1, 6}, {"u", 1, 20}};
  const int ndata = (nhydro > 0 ? 5 : 0) + (nstar > 0 ? 1 : 0);
  idata[19] = 0;                       // no unit names: dimensionless
  idata[20] = ndata;
  ilpdata[0] = (long) counts2[0]; ilpdata[1] = (long) counts2[1];
  rdata[0] = (float) h_fac;
  ddata[0] = times3[0]; ddata[1] = times3[1]; ddata[2] = times3[2];
  su_str(f, kSuTag);
  su_put(f, (int) sizeof(float));
  for (int k = 0; k < 3; k++) su_put(f, 3);
  for (int k = 0; k < 50; k++) su_put(f, idata[k]);
  for (int k = 0; k < 50; k++) su_put(f, ilpdata[k]);
  for (int k = 0; k < 50; k++) su_put(f, rdata[k]);
  for (int k = 0; k < 50; k++) su_put(f, ddata[k]);
  if (nhydro > 0) for (const Arr& a : hyd) su_str(f, a.id);
  if (nstar > 0) su_str(f, "sink_v1");
  if (nhydro > 0)
    for (const Arr& a : hyd) {
      const int td[5] = {a.w, 1, nhydro, 4, a.unit};
      for (int k = 0; k < 5; k++) su_put(f, td[k]);
    }
  if (nstar > 0) {
    const int td[5] = {1, 1, nstar, 7, 0};
    for (int k = 0; k < 5; k++) su_put(f, td[k]);
  }
  const size_t n = (size_t) nhydro;
  auto put = [&f](const double* a, size_t cnt) {
    for (size_t i = 0; i < cnt; i++) su_put(f, (float) a[i]);
  };
  if (nhydro > 0) { put(r, 3*n); put(m, n); put(h, n); put(rho, n); put(u, n); }
  if (nstar > 0) {
    const int head[6] = {2, 2, 0, kSinkLen, 0, 0};
    for (int k = 0; k < 6; k++) su_put(f, head[k]);
    for (int i = 0; i < nstar; i++) {
      const size_t j = n + (size_t) i;
      su_put(f, 1); su_put(f, 1); su_put(f, i + 1); su_put(f, 0);
      float sd[kSinkLen] = {};
      for (int k = 0; k < 3; k++) {
        sd[1 + k] = (float) r[3*j + k]; sd[4 + k] = (float) v[3*j + k];
      }
      sd[7] = (float) m[j]; sd[8] = (float) h[j];
      for (int k = 0; k < kSinkLen; k++) su_put(f, sd[k]);
    }
  }
  return f.good() ? GDF_OK : GDF_ERR_INVALID;
}

int gdf_write_snapshot(gdf_ctx* c, const char* path, const char* format) {
  if (!c || !path || !format || c->N <= 0) return GDF_ERR_INVALID;
  const std::string fmt(format);
  const bool lite = fmt == "slite" || fmt == "seren_lite";
  if (fmt != "column" && fmt != "su" && fmt != "sf" && !lite)
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
  int rc;
  if (fmt == "column") {
    rc = gdf_format_column(path, nh, ns, c->t, r.data(), v.data(), m.data(), h.data(),
                           rho.data(), u.data());
  } else {
    // the mean gas mass as SphSimulation.cpp:260-262 sums it, in particle order
    double mmean = 0.0;
    for (int i = 0; i < nh; i++) mmean += m[i];
    if (nh > 0) mmean /= (double) nh;
    const int64_t counts[3] = {c->snap_nout, c->nsteps, c->snap_nlite};
    const double times[4] = {c->t, c->snap_tlast, mmean, c->snap_tlitelast};
    const int* porig = c->snap_porig.size() == (size_t) nh ? c->snap_porig.data() : nullptr;
    if (lite)
      rc = gdf_format_slite(path, nh, ns, counts, times, c->phys.h_fac, r.data(), v.data(),
                            m.data(), h.data(), rho.data(), u.data());
    else
      rc = (fmt == "su" ? gdf_format_su : gdf_format_sf)(path, nh, ns, counts, times,
                                                        c->phys.h_fac, porig, r.data(), v.data(),
                                                        m.data(), h.data(), rho.data(), u.data());
  }
  return rc == GDF_OK ? rc : gdf::fail(c, rc, std::string("cannot write ") + path);
}

// What a Seren read keeps (SimulationIO.hpp:1363-1394): the time and step count always; the
// snapshot counters only on a restart (otherwise the constructor's zeros); the file's particle
// ids until the setup, which resets them unless restarting (SphSimulation.cpp:215-217)
extern "C++" template <typename I>
static void keep_header(gdf_ctx* c, const I* ilpdata, const double* ddata,
                        std::vector<int>&& porig) {
  c->t = ddata[0];
  c->nsteps = ilpdata[1];
  if (c->cfg.restart) {
    c->snap_nout = ilpdata[0]; c->snap_nlite = ilpdata[10];
    c->snap_tlast = ddata[1]; c->snap_tlitelast = ddata[10];
  } else {
    c->snap_nout = 0; c->snap_nlite = 0;
    c->snap_tlast = 0.0; c->snap_tlitelast = 0.0;
  }
  c->snap_porig = std::move(porig);
}

// ReadSerenUnformSnapshotFile (SimulationIO.hpp:1294-1540) into a device context: the header's
// step count and time are taken as the reference takes them, and its snapshot counters are kept
// for the next write; "temp" is skipped, other per-particle arrays are refused as there.
static int read_su(gdf_ctx* c, const char* path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, std::string("cannot read ") + path);
  if (su_name(f) != kSuTag) return gdf::fail(c, GDF_ERR_INVALID, "not a Seren binary snapshot");
  int w = 0, d[3] = {};
  if (!su_get(f, w) || w != (int) sizeof(double))
    return gdf::fail(c, GDF_ERR_INVALID, "snapshot FLOAT size is not 8 bytes");
  for (int k = 0; k < 3; k++)
    if (!su_get(f, d[k]) || d[k] != 3) return gdf::fail(c, GDF_ERR_INVALID, "snapshot is not 3D");
  int idata[50];
  long ilpdata[50];
  double rdata[50], ddata[50];
  bool ok = true;
  for (int k = 0; k < 50; k++) ok = ok && su_get(f, idata[k]);
  for (int k = 0; k < 50; k++) ok = ok && su_get(f, ilpdata[k]);
  for (int k = 0; k < 50; k++) ok = ok && su_get(f, rdata[k]);
  for (int k = 0; k < 50; k++) ok = ok && su_get(f, ddata[k]);
  const int nh = idata[0], ns = idata[1], nunit = idata[19], ndata = idata[20];
  if (!ok || nh < 0 || ns < 0 || (int64_t) nh + ns <= 0 || (int64_t) nh + ns > (1LL << 30) ||
      nunit < 0 || nunit > 50 || ndata < 0 || ndata > 50)
    return gdf::fail(c, GDF_ERR_INVALID, "bad Seren snapshot header");
  if (c->phys.nbody ? nh != 0 : ns != 0)
    return gdf::fail(c, GDF_ERR_INVALID, c->phys.nbody ? "N-body context: hydro particles in snapshot"
                                                       : "hydro context: star particles in snapshot");
  for (int k = 0; k < nunit; k++) su_name(f);
  std::vector<std::string> ids(ndata);
  for (int k = 0; k < ndata; k++) ids[k] = su_name(f);
  std::vector<int> td(5*(size_t) ndata);
  for (int& x : td) ok = ok && su_get(f, x);
  if (!ok) return gdf::fail(c, GDF_ERR_INVALID, "truncated Seren snapshot header");
  const size_t n = (size_t) (nh + ns), nhs = (size_t) nh;
  std::vector<double> r(3*n), v(3*n), m(n), h(n), rho(n, 0.0), u(n, 0.0);
  std::vector<int> porig;
  for (int j = 0; j < ndata && ok; j++) {
    const std::string& id = ids[j];
    if (id == "porig") { porig.resize(nhs); for (int& x : porig) ok = ok && su_get(f, x); }
    else if (id == "r") { for (size_t i = 0; i < 3*nhs; i++) ok = ok && su_get(f, r[i]); }
    else if (id == "m") { for (size_t i = 0; i < nhs; i++) ok = ok && su_get(f, m[i]); }
    else if (id == "h") { for (size_t i = 0; i < nhs; i++) ok = ok && su_get(f, h[i]); }
    else if (id == "v") { for (size_t i = 0; i < 3*nhs; i++) ok = ok && su_get(f, v[i]); }
    else if (id == "rho") { for (size_t i = 0; i < nhs; i++) ok = ok && su_get(f, rho[i]); }
    else if (id == "u") { for (size_t i = 0; i < nhs; i++) ok = ok && su_get(f, u[i]); }
    else if (id == "temp") { f.seekg((std::streamoff) (sizeof(double)*nhs), std::ios_base::cur); }
    else if (id == "sink_v1") {
      int head[6];
      for (int& x : head) ok = ok && su_get(f, x);
      if (!ok || head[0] != 2 || head[1] != 2 || head[2] != 0 || head[3] != kSinkLen)
        return gdf::fail(c, GDF_ERR_INVALID, "bad sink_v1 block");
      for (int i = 0; i < ns && ok; i++) {
        int flags[4];
        double sd[kSinkLen];
        for (int& x : flags) ok = ok && su_get(f, x);
        for (double& x : sd) ok = ok && su_get(f, x);
        const size_t k = nhs + (size_t) i;
        for (int a = 0; a < 3; a++) { r[3*k + a] = sd[1 + a]; v[3*k + a] = sd[4 + a]; }
        m[k] = sd[7]; h[k] = sd[8];
      }
    } else if (td[5*(size_t) j] >= 1) {
      return gdf::fail(c, GDF_ERR_INVALID, "snapshot array " + id + " cannot be read");
    }
  }
  if (!ok || !f) return gdf::fail(c, GDF_ERR_INVALID, "truncated Seren snapshot");
  int rc = gdf_set_particles(c, (int) n, r.data(), v.data(), m.data(), h.data(), u.data());
  if (rc != GDF_OK) return rc;
  if (!c->phys.nbody) rc = gdf_set_field(c, "rho", rho.data());
  if (rc != GDF_OK) return rc;
  keep_header(c, ilpdata, ddata, std::move(porig));
  (void) rdata;
  return GDF_OK;
}

// ReadSerenFormSnapshotFile (SimulationIO.hpp:601-925): the formatted twin of read_su.
static int read_sf(gdf_ctx* c, const char* path) {
  std::ifstream f(path);
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, std::string("cannot read ") + path);
  std::string tag;
  f >> tag;
  if (tag != "SERENASCIIDUMPV2" && tag != "SERENASCIIDUMPV3")
    return gdf::fail(c, GDF_ERR_INVALID, "not a Seren formatted snapshot");
  int w = 0, d[3] = {};
  f >> w >> d[0] >> d[1] >> d[2];
  if (!f || d[0] != 3 || d[1] != 3 || d[2] != 3)
    return gdf::fail(c, GDF_ERR_INVALID, "snapshot is not 3D");
  long long idata[50], ilpdata[50];
  double rdata[50], ddata[50];
  for (long long& x : idata) f >> x;
  for (long long& x : ilpdata) f >> x;
  for (double& x : rdata) f >> x;
  for (double& x : ddata) f >> x;
  const long long nh = idata[0], ns = idata[1], nunit = idata[19], ndata = idata[20];
  if (!f || nh < 0 || ns < 0 || (int64_t) nh + ns <= 0 || (int64_t) nh + ns > (1LL << 30) || nunit < 0 ||
      nunit > 50 || ndata < 0 || ndata > 50)
    return gdf::fail(c, GDF_ERR_INVALID, "bad Seren snapshot header");
  if (c->phys.nbody ? nh != 0 : ns != 0)
    return gdf::fail(c, GDF_ERR_INVALID, c->phys.nbody ? "N-body context: hydro particles in snapshot"
                                                       : "hydro context: star particles in snapshot");
  std::string tmp;
  for (long long k = 0; k < nunit; k++) f >> tmp;
  std::vector<std::string> ids((size_t) ndata);
  for (std::string& x : ids) f >> x;
  std::vector<long long> td(5*(size_t) ndata);
  for (long long& x : td) f >> x;
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, "truncated Seren snapshot header");
  const size_t n = (size_t) (nh + ns), nhs = (size_t) nh;
  std::vector<double> r(3*n), v(3*n), m(n), h(n), rho(n, 0.0), u(n, 0.0);
  std::vector<int> porig;
  for (size_t j = 0; j < ids.size() && f; j++) {
    const std::string& id = ids[j];
    if (id == "porig") { porig.resize(nhs); for (int& x : porig) f >> x; }
    else if (id == "r") { for (size_t i = 0; i < 3*nhs; i++) f >> r[i]; }
    else if (id == "m") { for (size_t i = 0; i < nhs; i++) f >> m[i]; }
    else if (id == "h") { for (size_t i = 0; i < nhs; i++) f >> h[i]; }
    else if (id == "v") { for (size_t i = 0; i < 3*nhs; i++) f >> v[i]; }
    else if (id == "rho") { for (size_t i = 0; i < nhs; i++) f >> rho[i]; }
    else if (id == "u") { for (size_t i = 0; i < nhs; i++) f >> u[i]; }
    else if (id == "temp") { double x; for (size_t i = 0; i < nhs; i++) f >> x; }
    else if (id == "sink_v1") {
      long long head[6];
      for (long long& x : head) f >> x;
      if (!f || head[0] != 2 || head[1] != 2 || head[2] != 0 || head[3] != kSinkLen)
        return gdf::fail(c, GDF_ERR_INVALID, "bad sink_v1 block");
      for (long long i = 0; i < ns && f; i++) {
        long long id1, zero;
        double sd[kSinkLen];
        f >> tmp >> id1 >> zero;
        for (double& x : sd) f >> x;
        const size_t k = nhs + (size_t) i;
        for (int a = 0; a < 3; a++) { r[3*k + a] = sd[1 + a]; v[3*k + a] = sd[4 + a]; }
        m[k] = sd[7]; h[k] = sd[8];
      }
    } else if (td[5*j] >= 1) {
      return gdf::fail(c, GDF_ERR_INVALID, "snapshot array " + id + " cannot be read");
    }
  }
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, "truncated Seren snapshot");
  int rc = gdf_set_particles(c, (int) n, r.data(), v.data(), m.data(), h.data(), u.data());
  if (rc != GDF_OK) return rc;
  if (!c->phys.nbody) rc = gdf_set_field(c, "rho", rho.data());
  if (rc != GDF_OK) return rc;
  keep_header(c, ilpdata, ddata, std::move(porig));
  (void) w; (void) rdata;
  return GDF_OK;
}

static int read_column(gdf_ctx* c, const char* path);

int gdf_read_snapshot(gdf_ctx* c, const char* path, const char* format) {
  if (!c || !path || !format) return GDF_ERR_INVALID;
  try {
    if (std::string(format) == "su") return read_su(c, path);
    if (std::string(format) == "sf") return read_sf(c, path);
    if (std::string(format) == "column") return read_column(c, path);
  } catch (const std::bad_alloc&) {
    return gdf::fail(c, GDF_ERR_NOMEM, "snapshot read: host allocation failed");
  }
  return gdf::fail(c, GDF_ERR_INVALID, std::string("unsupported snapshot format ") + format);
}

}  // extern "C"

// ReadColumnSnapshotFile (SimulationIO.hpp:193-270)
static int read_column(gdf_ctx* c, const char* path) {
  std::ifstream f(path);
  if (!f) return gdf::fail(c, GDF_ERR_INVALID, std::string("cannot read ") + path);
  int nh = 0, ns = 0, ndim = 0;
  double t = 0.0;
  f >> nh >> ns >> ndim >> t;
  if (!f || ndim != 3 || nh < 0 || ns < 0 || (int64_t) nh + ns <= 0 ||
      (int64_t) nh + ns > (1LL << 30))
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
  c->t = t;                         // ReadColumnSnapshotFile: t and tsnaplast from the header
  if (!c->cfg.restart) { c->snap_nout = 0; c->snap_nlite = 0; c->snap_tlitelast = 0.0; }
  c->snap_tlast = t;
  return rc;
}

extern "C" {

int gdf_snapshot_counters(gdf_ctx* c, const int64_t* set2, const double* set2d, int64_t* get2,
                          double* get2d) {
  if (!c) return GDF_ERR_INVALID;
  if (get2) { get2[0] = c->snap_nout; get2[1] = c->snap_nlite; }
  if (get2d) { get2d[0] = c->snap_tlast; get2d[1] = c->snap_tlitelast; }
  if (set2) {
    if (set2[0] < 0 || set2[1] < 0) return gdf::fail(c, GDF_ERR_INVALID, "negative snapshot count");
    c->snap_nout = set2[0]; c->snap_nlite = set2[1];
  }
  if (set2d) { c->snap_tlast = set2d[0]; c->snap_tlitelast = set2d[1]; }
  return GDF_OK;
}

}  // extern "C"
