// Ewald periodic-gravity correction field for a fully periodic box: the table of
// Ewald::Ewald (src/Common/Ewald.cpp:42-420, ewald_periodicity = 7) built on the host at
// context creation.  Each grid node x in [0, EFratio L/2]^3 (nEwaldGrid nodes per axis over
// nEwaldGrid - 2 cells of size EFratio (L/2)/(nEwaldGrid - 2), one spare node) holds the
// potential and acceleration of a unit mass and all its periodic images minus the nearest one:
//   pot = sum_n erfc(alpha |x + n L|)/|x + n L|
//         + sum_{h != 0} exp(-dzeta c3) cos(c4)/(pi L_x ratio1 ratio2 c3) - 1/|x|,
//   acc = -grad pot (short-range terms AccShort, long-range terms AccLong3pPer),
// with alpha = 2 ewald_mult/L_x, dzeta = (pi/(alpha L_x))^2, c3 = h_x^2 + (h_y/ratio1)^2 +
// (h_z/ratio2)^2, c4 = 2 pi (h_x x + h_y y/ratio1 + h_z z/ratio2)/L_x, the real-space images
// inside the ellipsoid cr.n^2 <= nseries^2 and the Fourier terms inside cf.h^2 <= nseries^2
// (the reference's truncation).  As in the reference, the k = 0 (neutralising background)
// constant of the potential is not added, and the node at the origin is set to zero.  The
// node layout here is interleaved (pot, ax, ay, az) per node for one 32-byte load per corner.
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "../gdf_internal.h"

namespace gdf {

int ewald_build_table(const double L[3], int nseries, int ngrid, double ewald_mult,
                      double efratio, std::vector<double>& field, int Ng[3], double dI[3]) {
  if (nseries < 1 || nseries > 64 || ngrid < 4 || ngrid > 256 || !(ewald_mult > 0.0) ||
      !(efratio >= 1.0) || !(efratio <= 1.5))
    return GDF_ERR_INVALID;
  // the reference's constants pi and twopi (Constants.h:61-62)
  const double pi = 3.14159265358979, twopi = 6.28318530717959;
  const double lx = L[0], ly = L[1], lz = L[2];
  const double linv = 1.0/lx;
  const double ratio1 = ly/lx, ratio2 = lz/lx;
  const double rinv1 = 1.0/ratio1, rinv2 = 1.0/ratio2;
  const int nrx = nseries, nry = (int) (1 + nseries/ratio1), nrz = (int) (1 + nseries/ratio2);
  const int nfx = nseries, nfy = (int) (1 + nseries*ratio1), nfz = (int) (1 + nseries*ratio2);
  const double cr2 = ratio1*ratio1, cr3 = ratio2*ratio2;
  const double cf2 = 1.0/(ratio1*ratio1), cf3 = 1.0/(ratio2*ratio2);
  const int radius2 = nseries*nseries;
  const double linv2 = linv*rinv1*rinv2/pi;
  const double linv_acc = 2.0*linv*linv*rinv1*rinv2;
  const double alpha = 2.0*linv*ewald_mult;
  const double dzeta = pi*pi*linv*linv/(alpha*alpha);
  const double sqrtpi = std::sqrt(pi);
  double Lew[3];
  int Nc[3];
  for (int k = 0; k < 3; k++) {
    Lew[k] = efratio*0.5*L[k];
    Nc[k] = ngrid - 1;
    dI[k] = (Nc[k] - 1)/Lew[k];
    Ng[k] = Nc[k] + 1;
  }
  const size_t one = (size_t) Ng[0]*Ng[1]*Ng[2];
  field.assign(4*one, 0.0);
  // the Fourier terms of the truncated set, shared by every node
  struct Hterm { int hx, hy, hz; double c3, w; };
  std::vector<Hterm> hs;
  for (int hx = -nfx; hx <= nfx; hx++)
    for (int hy = -nfy; hy <= nfy; hy++)
      for (int hz = -nfz; hz <= nfz; hz++) {
        if (!(hx*hx + cf2*hy*hy + cf3*hz*hz <= radius2)) continue;
        if (hx*hx + hy*hy + hz*hz == 0) continue;
        const double c3 = hx*hx + (hy*rinv1)*(hy*rinv1) + (hz*rinv2)*(hz*rinv2);
        hs.push_back({hx, hy, hz, c3, std::exp(-dzeta*c3)/c3});
      }
  struct Nterm { int nx, ny, nz; };
  std::vector<Nterm> ns;
  for (int nx = -nrx; nx <= nrx; nx++)
    for (int ny = -nry; ny <= nry; ny++)
      for (int nz = -nrz; nz <= nrz; nz++)
        if (nx*nx + cr2*ny*ny + cr3*nz*nz <= radius2) ns.push_back({nx, ny, nz});
  auto node = [&](int i, int j, int k) {
    const double x = i*Lew[0]/(Nc[0] - 1), y = j*Lew[1]/(Nc[1] - 1), z = k*Lew[2]/(Nc[2] - 1);
    double pot = 0.0, acc[3] = {0.0, 0.0, 0.0};
    for (const Nterm& n : ns) {
      const double rr[3] = {x + n.nx*lx, y + n.ny*ly, z + n.nz*lz};
      const double r2 = rr[0]*rr[0] + rr[1]*rr[1] + rr[2]*rr[2];
      const double r = std::sqrt(r2);
      const double ec = std::erfc(alpha*r);
      pot += ec/r;
      const double g = 2.0*alpha*std::exp(-r2*alpha*alpha)/(sqrtpi*r2) + ec/(r2*r);
      for (int d = 0; d < 3; d++) acc[d] += rr[d]*g;
    }
    for (const Hterm& h : hs) {
      const double c4 = twopi*linv*(h.hx*x + h.hy*y*rinv1 + h.hz*z*rinv2);
      pot += linv2*h.w*std::cos(c4);
      const double s = linv_acc*h.w*std::sin(c4);
      acc[0] += h.hx*s;
      acc[1] += rinv1*(h.hy*s);
      acc[2] += rinv2*(h.hz*s);
    }
    const double r = std::sqrt(x*x + y*y + z*z);
    const double r3 = r*r*r;
    const size_t e = 4*((size_t) k*Ng[1]*Ng[0] + (size_t) j*Ng[0] + i);
    field[e] = pot - 1.0/r;
    field[e + 1] = acc[0] - x/r3;
    field[e + 2] = acc[1] - y/r3;
    field[e + 3] = acc[2] - z/r3;
  };
  const int nth = (int) std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  for (int t = 0; t < nth; t++)
    pool.emplace_back([&, t]() {
      for (int k = t; k < Ng[2]; k += nth)
        for (int j = 0; j < Ng[1]; j++)
          for (int i = 0; i < Ng[0]; i++) node(i, j, k);
    });
  for (std::thread& th : pool) th.join();
  for (int c = 0; c < 4; c++) field[c] = 0.0;     // the origin's 1/0 terms
  return GDF_OK;
}

}  // namespace gdf
