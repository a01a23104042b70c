// Stars beside the gas: the N-body half of SphSimulation::MainLoop
// (src/Hydrodynamics/SphSimulation.cpp:595, 775-833, 881) on the device.
//
//   * gas from stars: GradhSph::ComputeStarGravForces (src/GradhSph/GradhSph.cpp:699-743), the
//     kernel-softened pull of every star on every active gas particle, over the mean smoothing
//     length 2/(h_i + h_s);
//   * stars from gas: HydroTree::UpdateAllStarGasForces (src/Tree/HydroTree.cpp:552-654) with
//     the sums of NbodyLeapfrogKDK::CalculateDirectHydroForces (src/Nbody/NbodyLeapfrogKDK.cpp:
//     151-240): a gas particle within k (h_s + h_j)/2 of the star acts through the softened
//     kernel, any other one as a point mass.  The reference reaches distant gas through tree
//     cells; here every gas particle is summed directly (its cell-free limit: the reference's
//     brute-force tree gives the same sums), since a run holds few stars and the pass is one
//     stream over the gas;
//   * the Hubber et al. (2013) conservative SPH-star zeta terms of the h-rho pass
//     (GradhSph.cpp:285-307) are added in the density pass (gdf_sph.hip);
//   * star-star sums: the N-body direct sums of gdf_gravity.hip on the star arrays;
//   * the stars' leapfrog KDK: NbodyLeapfrogKDK::AdvanceParticles, CorrectionTerms,
//     EndTimestep and Timestep (src/Nbody/NbodyLeapfrogKDK.cpp:253-401).
// All in FP64 whatever the gas pair precision: the stars are few and their sums are direct.
// Star-gas separations are the plain coordinate differences in both directions: the reference's
// star-from-gas sums form a periodic image into a vector they then do not use
// (NbodyLeapfrogKDK.cpp:183-186) and its gas-from-star sums form none.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;

struct StarView {
  int n;
  const double* x; const double* y; const double* z; const double* m; const double* h;
};

StarView star_view(const gdf_ctx* c) {
  StarView s;
  s.n = c->Nstar;
  s.x = c->sv[S_X]; s.y = c->sv[S_Y]; s.z = c->sv[S_Z]; s.m = c->sv[S_M]; s.h = c->sv[S_H];
  return s;
}

struct GasFromStarsArgs {
  int N;
  const int* act;
  const double* x; const double* y; const double* z; const double* h;
  double* atx; double* aty; double* atz; double* ax; double* ay; double* az; double* gpot;
  StarView s;
  PhysParams p;
};

// GradhSph::ComputeStarGravForces per active gas particle, stars staged in LDS
__global__ __launch_bounds__(BS) void k_gas_from_stars(GasFromStarsArgs a) {
#pragma clang fp contract(off)
  __shared__ double sx[BS], sy[BS], sz[BS], sm[BS], sh[BS];
  const int i = blockIdx.x*BS + threadIdx.x;
  const bool has = i < a.N && (!a.act || a.act[i]);
  double xi = 0.0, yi = 0.0, zi = 0.0, hi = 1.0;
  if (has) { xi = a.x[i]; yi = a.y[i]; zi = a.z[i]; hi = a.h[i]; }
  double gx = 0.0, gy = 0.0, gz = 0.0, gp = 0.0;
  const KernelDesc& K = a.p.kern;
  for (int j0 = 0; j0 < a.s.n; j0 += BS) {
    const int jl = j0 + threadIdx.x;
    if (jl < a.s.n) {
      sx[threadIdx.x] = a.s.x[jl]; sy[threadIdx.x] = a.s.y[jl]; sz[threadIdx.x] = a.s.z[jl];
      sm[threadIdx.x] = a.s.m[jl]; sh[threadIdx.x] = a.s.h[jl];
    }
    __syncthreads();
    const int nb = min(BS, a.s.n - j0);
    if (has) {
      for (int q = 0; q < nb; q++) {
        const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
        const double drsqd = dx*dx + dy*dy + dz*dz + 1.0e-20;
        const double drmag = sqrt(drsqd);
        const double invdrmag = 1.0/drmag;
        const double invhmean = 2.0/(hi + sh[q]);
        const double ms = sm[q];
        const double paux = ms*invhmean*invhmean*kwgrav<double>(K, drmag*invhmean)*invdrmag;
        gx += paux*dx; gy += paux*dy; gz += paux*dz;
        gp += ms*invhmean*kwpot<double>(K, drmag*invhmean);
      }
    }
    __syncthreads();
  }
  if (!has) return;
  a.atx[i] += gx; a.aty[i] += gy; a.atz[i] += gz;
  a.ax[i] += gx; a.ay[i] += gy; a.az[i] += gz;
  a.gpot[i] += gp;
}

// x^3 rounded once (to within the last bit), as the reference's pow(x, 3): x^2 split exactly into
// a rounded product and its error, each multiplied by x with the first product's error kept
__device__ __forceinline__ double cube_rn(double x) {
  const double s = x*x, se = fma(x, x, -s);
  const double c = s*x, ce = fma(s, x, -c);
  return c + (ce + se*x);
}

struct StarsFromGasArgs {
  int N, nchunk;
  const double* x; const double* y; const double* z; const double* m; const double* h;
  StarView s;
  double* part;     // [star][chunk][4]: a_x, a_y, a_z, gpot
  PhysParams p;
};

// One workgroup per (chunk of kStarChunk gas particles, star): the star's sums over its chunk,
// reduced in a fixed order (no atomics: the result does not depend on the schedule)
__global__ __launch_bounds__(BS) void k_stars_from_gas(StarsFromGasArgs a) {
#pragma clang fp contract(off)
  __shared__ double red[4][BS];
  const int star = blockIdx.y, chunk = blockIdx.x;
  const double xs = a.s.x[star], ys = a.s.y[star], zs = a.s.z[star], hs = a.s.h[star];
  const KernelDesc& K = a.p.kern;
  const double kr = K.range;
  double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
  const int j1 = min(a.N, (chunk + 1)*kStarChunk);
  for (int j = chunk*kStarChunk + threadIdx.x; j < j1; j += BS) {
    const double dx = a.x[j] - xs, dy = a.y[j] - ys, dz = a.z[j] - zs;
    const double drsqd = dx*dx + dy*dy + dz*dz;
    const double mj = a.m[j], hj = a.h[j];
    const double rn = 0.5*(kr*hs) + (0.5*kr)*hj;
    const double drmag = sqrt(drsqd);
    const double invdrmag = 1.0/drmag;
    double paux;
    if (drsqd < rn*rn) {
      const double invhmean = 2.0/(hs + hj);
      paux = mj*invhmean*invhmean*kwgrav<double>(K, drmag*invhmean)*invdrmag;
      pot += mj*invhmean*kwpot<double>(K, drmag*invhmean);
    } else {
      paux = mj*cube_rn(invdrmag);
      pot += mj*invdrmag;
    }
    ax += paux*dx; ay += paux*dy; az += paux*dz;
  }
  red[0][threadIdx.x] = ax; red[1][threadIdx.x] = ay; red[2][threadIdx.x] = az;
  red[3][threadIdx.x] = pot;
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 4; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) a.part[((size_t) star*a.nchunk + chunk)*4 + threadIdx.x] = red[threadIdx.x][0];
}

// per star, the chunks' partial sums in chunk order, written over a and gpot with adot cleared
// (the MainLoop clears the star terms before this pass, SphSimulation.cpp:778-787), or added
__global__ void k_stars_reduce(int ns, int nchunk, int add, const double* __restrict__ part,
                               double* ax, double* ay, double* az, double* adx, double* ady,
                               double* adz, double* gpot) {
  const int s = blockIdx.x*BS + threadIdx.x;
  if (s >= ns) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int c = 0; c < nchunk; c++)
    for (int k = 0; k < 4; k++) acc[k] += part[((size_t) s*nchunk + c)*4 + k];
  if (add) {
    ax[s] += acc[0]; ay[s] += acc[1]; az[s] += acc[2]; gpot[s] += acc[3];
    return;
  }
  ax[s] = acc[0]; ay[s] = acc[1]; az[s] = acc[2]; gpot[s] = acc[3];
  adx[s] = 0.0; ady[s] = 0.0; adz[s] = 0.0;
}

struct StarKdkArgs {
  int n;
  double dt;
  double* r[3]; double* v[3]; double* a[3]; double* r0[3]; double* v0[3]; double* a0[3];
};

// NbodyLeapfrogKDK::AdvanceParticles (NbodyLeapfrogKDK.cpp:253-289), dt = t - tlast
__global__ void k_stars_advance(StarKdkArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.n) return;
  for (int k = 0; k < 3; k++) {
    a.r[k][i] = a.r0[k][i] + a.v0[k][i]*a.dt + 0.5*a.a0[k][i]*a.dt*a.dt;
    a.v[k][i] = a.v0[k][i] + a.a0[k][i]*a.dt;
  }
}

// NbodyLeapfrogKDK::CorrectionTerms (:296-326): v += (a - a0)(t - tlast)/2
__global__ void k_stars_correction(StarKdkArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.n) return;
  for (int k = 0; k < 3; k++) a.v[k][i] += 0.5*(a.a[k][i] - a.a0[k][i])*a.dt;
}

// NbodyLeapfrogKDK::EndTimestep (:333-370)
__global__ void k_stars_end(StarKdkArgs a) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.n) return;
  for (int k = 0; k < 3; k++) {
    a.r0[k][i] = a.r[k][i]; a.v0[k][i] = a.v[k][i]; a.a0[k][i] = a.a[k][i];
  }
}

// NbodyLeapfrogKDK::Timestep (:386-401) minimum over the stars, and the stars' diagnostics
// sums (SimAnalysis.hpp:106-122) in star order; one workgroup
__global__ void k_stars_scalars(int n, double mult, const double* x, const double* y,
                                const double* z, const double* vx, const double* vy,
                                const double* vz, const double* ax, const double* ay,
                                const double* az, const double* m, const double* h,
                                const double* gpot, double* out) {
#pragma clang fp contract(off)
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double dtmin = 9.9e50;
  double acc[11] = {};
  for (int i = 0; i < n; i++) {
    const double amag = sqrt(ax[i]*ax[i] + ay[i]*ay[i] + az[i]*az[i]);
    dtmin = fmin(dtmin, mult*sqrt(h[i]/(amag + 1.0e-50)));
    const double mi = m[i];
    acc[0] += mi;
    acc[1] += mi*(vx[i]*vx[i] + vy[i]*vy[i] + vz[i]*vz[i]);
    acc[3] -= mi*gpot[i];
    acc[4] += mi*vx[i]; acc[5] += mi*vy[i]; acc[6] += mi*vz[i];
    acc[7] += mi*(y[i]*vz[i] - z[i]*vy[i]);
    acc[8] += mi*(z[i]*vx[i] - x[i]*vz[i]);
    acc[9] += mi*(x[i]*vy[i] - y[i]*vx[i]);
  }
  out[0] = dtmin;
  for (int k = 0; k < 11; k++) out[1 + k] = acc[k];
}

StarKdkArgs kdk_args(gdf_ctx* c, double dt) {
  StarKdkArgs a;
  a.n = c->Nstar; a.dt = dt;
  for (int k = 0; k < 3; k++) {
    a.r[k] = c->sv[S_X + k]; a.v[k] = c->sv[S_VX + k]; a.a[k] = c->sv[S_AX + k];
    a.r0[k] = c->sv[S_X0 + k]; a.v0[k] = c->sv[S_VX0 + k]; a.a0[k] = c->sv[S_AX0 + k];
  }
  return a;
}

// the timestep minimum and diagnostics sums of the stars into out[0, 12)
int stars_scalars(gdf_ctx* c, double* out) {
  double* d = c->red + 16384 - 16;      // the last 16 reduction slots
  k_stars_scalars<<<1, 64, 0, c->stream>>>(c->Nstar, c->phys.nbody_mult, c->sv[S_X],
                                           c->sv[S_Y], c->sv[S_Z], c->sv[S_VX], c->sv[S_VY],
                                           c->sv[S_VZ], c->sv[S_AX], c->sv[S_AY], c->sv[S_AZ],
                                           c->sv[S_M], c->sv[S_H], c->sv[S_GPOT], d);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipMemcpyAsync(c->h_red + 16384 - 16, d, sizeof(double)*12, hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  std::memcpy(out, c->h_red + 16384 - 16, sizeof(double)*12);
  return GDF_OK;
}

int star_field(const char* name, int* f, int* w) {
  struct E { const char* n; int f, w; };
  static const E tab[] = {{"r", S_X, 3}, {"v", S_VX, 3}, {"a", S_AX, 3}, {"adot", S_ADX, 3},
                          {"r0", S_X0, 3}, {"v0", S_VX0, 3}, {"a0", S_AX0, 3}, {"m", S_M, 1},
                          {"h", S_H, 1}, {"gpot", S_GPOT, 1}};
  for (const E& e : tab)
    if (!strcmp(e.n, name)) { *f = e.f; *w = e.w; return 1; }
  return 0;
}

}  // namespace

void stars_free(gdf_ctx* c) {
  for (int f = 0; f < S_N; f++) { if (c->sv[f]) (void) hipFree(c->sv[f]); c->sv[f] = nullptr; }
  if (c->star_part) (void) hipFree(c->star_part);
  c->star_part = nullptr; c->star_part_cap = 0;
  c->Nstar = c->Nstar_cap = 0;
}

int stars_on_gas(gdf_ctx* c) {
  if (c->Nstar <= 0 || c->N <= 0) return GDF_OK;
  GasFromStarsArgs a;
  a.N = c->N;
  a.act = c->nlevels > 1 ? c->pi[I_ACT] : nullptr;
  a.x = c->pf[F_X]; a.y = c->pf[F_Y]; a.z = c->pf[F_Z]; a.h = c->pf[F_H];
  a.atx = c->df[D_ATX]; a.aty = c->df[D_ATY]; a.atz = c->df[D_ATZ];
  a.ax = c->df[D_AX]; a.ay = c->df[D_AY]; a.az = c->df[D_AZ]; a.gpot = c->df[D_GPOT];
  a.s = star_view(c);
  a.p = c->phys;
  k_gas_from_stars<<<nblocks(c->N, BS), BS, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int stars_gas_on_stars(gdf_ctx* c, int with_gas, int add) {
  if (c->Nstar <= 0 || (add && !with_gas)) return GDF_OK;
  const int nchunk = with_gas ? std::max(1, nblocks(c->N, kStarChunk)) : 0;
  if (nchunk > 0) {
    const size_t need = (size_t) c->Nstar*nchunk*4;
    if (need > c->star_part_cap) {
      if (c->star_part) (void) hipFree(c->star_part);
      c->star_part = nullptr; c->star_part_cap = 0;
      GDF_HIP(c, hipMalloc(&c->star_part, sizeof(double)*need));
      c->star_part_cap = need;
    }
    StarsFromGasArgs a;
    a.N = c->N; a.nchunk = nchunk;
    a.x = c->pf[F_X]; a.y = c->pf[F_Y]; a.z = c->pf[F_Z]; a.m = c->pf[F_M]; a.h = c->pf[F_H];
    a.s = star_view(c);
    a.part = c->star_part;
    a.p = c->phys;
    k_stars_from_gas<<<dim3(nchunk, c->Nstar), BS, 0, c->stream>>>(a);
    GDF_HIP(c, hipGetLastError());
  }
  k_stars_reduce<<<nblocks(c->Nstar, BS), BS, 0, c->stream>>>(
      c->Nstar, nchunk, add, c->star_part, c->sv[S_AX], c->sv[S_AY], c->sv[S_AZ], c->sv[S_ADX],
      c->sv[S_ADY], c->sv[S_ADZ], c->sv[S_GPOT]);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int stars_advance(gdf_ctx* c, double t) {
  if (c->Nstar <= 0) return GDF_OK;
  k_stars_advance<<<nblocks(c->Nstar, BS), BS, 0, c->stream>>>(kdk_args(c, t - c->star_tlast));
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int stars_correction(gdf_ctx* c, double t) {
  if (c->Nstar <= 0) return GDF_OK;
  k_stars_correction<<<nblocks(c->Nstar, BS), BS, 0, c->stream>>>(
      kdk_args(c, t - c->star_tlast));
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int stars_end_step(gdf_ctx* c, double t) {
  if (c->Nstar <= 0) return GDF_OK;
  k_stars_end<<<nblocks(c->Nstar, BS), BS, 0, c->stream>>>(kdk_args(c, 0.0));
  GDF_HIP(c, hipGetLastError());
  c->star_tlast = t;
  return GDF_OK;
}

int stars_timestep(gdf_ctx* c, double* dtmin) {
  if (c->Nstar <= 0) return GDF_OK;
  double out[12];
  GDF_TRY(stars_scalars(c, out));
  if (!(out[0] > 0.0)) return fail(c, GDF_ERR_TIMESTEP, "Error : N-body timestep invalid");
  *dtmin = std::min(*dtmin, out[0]);
  return GDF_OK;
}

int stars_diagnostics(gdf_ctx* c, double* acc) {
  if (c->Nstar <= 0) return GDF_OK;
  double out[12];
  GDF_TRY(stars_scalars(c, out));
  for (int k = 0; k < 11; k++) acc[k] += out[1 + k];
  return GDF_OK;
}

}  // namespace gdf

using namespace gdf;

extern "C" {

int gdf_set_stars(gdf_ctx* c, int nstar, const double* r, const double* v, const double* m,
                  const double* h) {
  if (!c || nstar < 0 || (nstar > 0 && (!r || !v || !m || !h)))
    return c ? fail(c, GDF_ERR_INVALID, "gdf_set_stars: bad arguments") : GDF_ERR_INVALID;
  if (c->phys.nbody) return fail(c, GDF_ERR_INVALID, "gdf_set_stars: an N-body context holds "
                                                     "its stars as its particles");
  if (c->phys.ewald && nstar > 0)
    return fail(c, GDF_ERR_INVALID, "stars beside the gas: periodic gravity is not built");
  if (nstar > (1 << 20)) return fail(c, GDF_ERR_INVALID, "gdf_set_stars: too many stars");
  for (int i = 0; i < nstar; i++)
    if (!(m[i] > 0.0) || !(h[i] > 0.0))
      return fail(c, GDF_ERR_INVALID, "star masses and smoothing lengths must be positive");
  if (nstar > c->Nstar_cap) {
    stars_free(c);
    for (int f = 0; f < S_N; f++) GDF_HIP(c, hipMalloc(&c->sv[f], sizeof(double)*nstar));
    c->Nstar_cap = nstar;
  }
  c->Nstar = nstar;
  c->star_tlast = c->t;
  if (nstar == 0) return GDF_OK;
  std::vector<double> col(nstar);
  auto up = [&](int f, const double* src, int stride, int off) -> int {
    for (int i = 0; i < nstar; i++) col[i] = src ? src[(size_t) i*stride + off] : 0.0;
    GDF_HIP(c, hipMemcpy(c->sv[f], col.data(), sizeof(double)*nstar, hipMemcpyHostToDevice));
    return GDF_OK;
  };
  for (int k = 0; k < 3; k++) {
    GDF_TRY(up(S_X + k, r, 3, k)); GDF_TRY(up(S_X0 + k, r, 3, k));
    GDF_TRY(up(S_VX + k, v, 3, k)); GDF_TRY(up(S_VX0 + k, v, 3, k));
    for (int f : {S_AX + k, S_ADX + k, S_AX0 + k}) GDF_TRY(up(f, nullptr, 1, 0));
  }
  GDF_TRY(up(S_M, m, 1, 0));
  GDF_TRY(up(S_H, h, 1, 0));
  GDF_TRY(up(S_GPOT, nullptr, 1, 0));
  c->setup_done = false;
  return GDF_OK;
}

int gdf_num_stars(const gdf_ctx* c) { return c ? c->Nstar : 0; }

int gdf_get_star_field(gdf_ctx* c, const char* name, double* out) {
  if (!c || !name || !out) return GDF_ERR_INVALID;
  int f = 0, w = 0;
  if (!star_field(name, &f, &w))
    return fail(c, GDF_ERR_INVALID, std::string("unknown star field ") + name);
  const int n = c->Nstar;
  if (n == 0) return GDF_OK;
  std::vector<double> col(n);
  for (int k = 0; k < w; k++) {
    GDF_HIP(c, hipMemcpy(col.data(), c->sv[f + k], sizeof(double)*n, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) out[(size_t) i*w + k] = col[i];
  }
  return GDF_OK;
}

int gdf_update_all_star_gas_forces(gdf_ctx* c) {
  if (!c) return GDF_ERR_INVALID;
  if (c->N <= 0) return GDF_OK;
  GDF_TRY(stars_gas_on_stars(c, 1, 1));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_nbody_star_forces(gdf_ctx* c, int zero_first) {
  if (!c) return GDF_ERR_INVALID;
  if (zero_first) GDF_TRY(stars_gas_on_stars(c, 0, 0));
  else GDF_TRY(nbody_direct_stars(c, c->cfg.nbody_softening));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_nbody_advance_particles(gdf_ctx* c, double t) {
  if (!c) return GDF_ERR_INVALID;
  GDF_TRY(stars_advance(c, t));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_nbody_correction_terms(gdf_ctx* c, double t) {
  if (!c) return GDF_ERR_INVALID;
  GDF_TRY(stars_correction(c, t));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_nbody_end_timestep(gdf_ctx* c, double t) {
  if (!c) return GDF_ERR_INVALID;
  GDF_TRY(stars_end_step(c, t));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

}  // extern "C"
