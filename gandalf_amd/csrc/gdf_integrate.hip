// Leapfrog KDK integrator, periodic wrap, timestep and diagnostics reductions.
//
// SphLeapfrogKDK::AdvanceParticles / EndTimestep (src/Hydrodynamics/SphLeapfrogKDK.cpp:76-127,
// :219-272), TimeIntegration::CheckBoundaries (src/Common/Integration.cpp:11-80),
// SphIntegration::Timestep (src/Hydrodynamics/SphIntegration.cpp:81-134) with the global minimum
// of Simulation::ComputeGlobalTimestep (src/Common/Simulation.cpp:1669-1754), and
// Simulation::CalculateDiagnostics (src/Common/SimAnalysis.hpp:52-185).  These streaming kernels
// are HBM-bound; integration arithmetic keeps the reference's operation order (no contraction)
// so positions and velocities are reproduced bit for bit from identical inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;
constexpr int RED_BLOCKS = 1024;

struct AdvArgs {
  int N, energy;
  double dt;
  const double* segdt; const int* seg;    // serving batches: the step of the particle's segment
  const double* r0[3]; const double* v0[3]; const double* a0[3];
  double* r[3]; double* v[3];
  const double* u0; const double* dudt0; double* u;
};

__global__ void k_advance(AdvArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.N) return;
  const double dt = a.segdt ? a.segdt[a.seg[i]] : a.dt;
  for (int k = 0; k < 3; k++) a.r[k][i] = a.r0[k][i] + a.v0[k][i]*dt + 0.5*a.a0[k][i]*dt*dt;
  for (int k = 0; k < 3; k++) a.v[k][i] = a.v0[k][i] + a.a0[k][i]*dt;
  if (a.energy) a.u[i] = a.u0[i] + a.dudt0[i]*dt;
}

struct WrapArgs { int N; double lo[3], hi[3], size[3]; double* r[3]; double* r0[3]; };

__global__ void k_wrap(WrapArgs a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.N) return;
  for (int k = 0; k < 3; k++) {
    double r = a.r[k][i];
    if (r < a.lo[k]) { a.r[k][i] = r + a.size[k]; a.r0[k][i] += a.size[k]; }
    r = a.r[k][i];
    if (r > a.hi[k]) { a.r[k][i] = r - a.size[k]; a.r0[k][i] -= a.size[k]; }
  }
}

struct EndArgs {
  int N, energy;
  double dt;
  const double* segdt; const int* seg;
  const double* a[3]; double* v[3]; double* r0[3]; double* v0[3]; double* a0[3];
  const double* r[3];
  const double* dudt; double* u; double* u0; double* dudt0;
};

__global__ void k_end_step(EndArgs e) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= e.N) return;
  const double dt = e.segdt ? e.segdt[e.seg[i]] : e.dt;
  for (int k = 0; k < 3; k++) {
    const double a = e.a[k][i], a0 = e.a0[k][i];
    const double v = e.v[k][i] + 0.5*dt*(a - a0);
    e.v[k][i] = v;
    e.r0[k][i] = e.r[k][i];
    e.v0[k][i] = v;
    e.a0[k][i] = a;
  }
  if (e.energy) {
    const double dudt = e.dudt[i], dudt0 = e.dudt0[i];
    double u = e.u[i] + 0.5*(dudt - dudt0)*dt;
    if (u <= 0.0) u = e.u0[i] + dudt0*dt;
    e.u[i] = u;
    e.u0[i] = u;
    e.dudt0[i] = dudt;
  }
}

struct TsArgs {
  int N, hydro, energy, nbody;
  double courant, accel, emult;
  const double* h; const double* sound; const double* div_v; const double* a[3];
  const double* u; const double* dudt;
  double* dt_out; double* part;
};

__global__ void k_timestep(TsArgs t) {
#pragma clang fp contract(off)
  __shared__ double s[BS];
  double best = 9.9e50;
  for (int i = blockIdx.x*BS + threadIdx.x; i < t.N; i += gridDim.x*BS) {
    const double h = t.h[i];
    double ts;
    if (t.nbody) {
      // NbodyLeapfrogKDK::Timestep (src/Nbody/NbodyLeapfrogKDK.cpp:386-401)
      const double ax = t.a[0][i], ay = t.a[1][i], az = t.a[2][i];
      ts = t.accel*sqrt(h/(sqrt(ax*ax + ay*ay + az*az) + 1.0e-50));
      if (t.dt_out) t.dt_out[i] = ts;
      best = fmin(best, ts);
      continue;
    }
    if (t.hydro) ts = t.courant*h/(t.sound[i] + h*fabs(t.div_v[i]) + 1.0e-50);
    else ts = t.courant*h/(h*fabs(t.div_v[i]) + 1.0e-50);
    const double ax = t.a[0][i], ay = t.a[1][i], az = t.a[2][i];
    const double amag = sqrt(ax*ax + ay*ay + az*az);
    ts = fmin(ts, t.accel*sqrt(h/(amag + 1.0e-50)));
    if (t.energy) ts = fmin(ts, t.emult*(t.u[i]/(fabs(t.dudt[i]) + 1.0e-20)));
    if (t.dt_out) t.dt_out[i] = ts;
    best = fmin(best, ts);
  }
  s[threadIdx.x] = best;
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) t.part[blockIdx.x] = s[0];
}

__global__ void k_min_final(int nb, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s[BS];
  double best = 9.9e50;
  for (int b = threadIdx.x; b < nb; b += BS) best = fmin(best, part[b]);
  s[threadIdx.x] = best;
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// per-segment minimum of the per-particle timesteps: one block per segment
__global__ void k_seg_min(const int* __restrict__ off, const double* __restrict__ v,
                          double* __restrict__ out) {
  __shared__ double sh[BS];
  const int sg = blockIdx.x;
  double best = 9.9e50;
  for (int i = off[sg] + threadIdx.x; i < off[sg + 1]; i += BS) best = fmin(best, v[i]);
  sh[threadIdx.x] = best;
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmin(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[sg] = sh[0];
}

__global__ void k_max_h(int N, const double* __restrict__ h, double* __restrict__ part) {
  __shared__ double s[BS];
  double best = 0.0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < N; i += gridDim.x*BS) best = fmax(best, h[i]);
  s[threadIdx.x] = best;
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

constexpr int ND = 11;  // mtot ke u gpe mom[3] ang[3] (+pad)

// stride: elements between one particle's r (and v) component and the next particle's (1 for
// the context's per-component arrays, 3 for (N,3) row-major arrays)
struct DiagArgs {
  int N, stride;
  const double* r[3]; const double* v[3]; const double* m; const double* u; const double* gpot;
  double* part;
};

__global__ void k_diag(DiagArgs d) {
  __shared__ double s[ND][BS];
  double acc[ND];
  for (int k = 0; k < ND; k++) acc[k] = 0.0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < d.N; i += gridDim.x*BS) {
    const double m = d.m[i];
    const size_t o = (size_t) i*d.stride;
    const double rx = d.r[0][o], ry = d.r[1][o], rz = d.r[2][o];
    const double vx = d.v[0][o], vy = d.v[1][o], vz = d.v[2][o];
    acc[0] += m;
    acc[1] += m*(vx*vx + vy*vy + vz*vz);
    acc[2] += m*d.u[i];
    acc[3] -= m*d.gpot[i];
    acc[4] += m*vx; acc[5] += m*vy; acc[6] += m*vz;
    acc[7] += m*(ry*vz - rz*vy);
    acc[8] += m*(rz*vx - rx*vz);
    acc[9] += m*(rx*vy - ry*vx);
  }
  for (int k = 0; k < ND; k++) s[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < ND; k++) s[k][threadIdx.x] += s[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < ND) d.part[blockIdx.x*ND + threadIdx.x] = s[threadIdx.x][0];
}

// The block partials of k_diag summed in block order, then the diagnostics as
// integ_diagnostics forms them on the host: out = {Etot, ketot, utot, gpetot, mom[3],
// angmom[3], mtot}.  One block of 64 threads.
__global__ void k_diag_final(int nb, const double* __restrict__ part, int hydro, int grav,
                             double* __restrict__ out) {
  __shared__ double acc[ND];
  const int k = threadIdx.x;
  if (k < ND) {
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[b*ND + k];
    acc[k] = s;
  }
  __syncthreads();
  if (k != 0) return;
  const double ketot = 0.5*acc[1], gpetot = 0.5*acc[3];
  double Etot = ketot;
  if (hydro) Etot += acc[2];
  if (grav) Etot += gpetot;
  out[0] = Etot; out[1] = ketot; out[2] = acc[2]; out[3] = gpetot;
  for (int j = 0; j < 3; j++) { out[4 + j] = acc[4 + j]; out[7 + j] = acc[7 + j]; }
  out[10] = acc[0];
}

__global__ void k_zero_d(int n, double* __restrict__ p) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

__global__ void k_fill_i(int n, int* __restrict__ p, int v) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_add_atree(int N, double* __restrict__ ax, double* __restrict__ ay,
                            double* __restrict__ az, const double* __restrict__ tx,
                            const double* __restrict__ ty, const double* __restrict__ tz,
                            const int* __restrict__ act) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N || (act && !act[i])) return;
  ax[i] += tx[i]; ay[i] += ty[i]; az[i] += tz[i];
}

__global__ void k_pack_posm(int N, const double* __restrict__ x, const double* __restrict__ y,
                            const double* __restrict__ z, const double* __restrict__ m,
                            double4* __restrict__ out) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < N) out[i] = make_double4(x[i], y[i], z[i], m[i]);
}

}  // namespace

int integ_advance(gdf_ctx* c, double dt) {
  AdvArgs a;
  a.N = c->N; a.energy = c->phys.energy_integration; a.dt = dt;
  for (int k = 0; k < 3; k++) {
    a.r0[k] = c->pf[F_X0 + k]; a.v0[k] = c->pf[F_VX0 + k]; a.a0[k] = c->pf[F_AX0 + k];
    a.r[k] = c->pf[F_X + k]; a.v[k] = c->pf[F_VX + k];
  }
  a.u0 = c->pf[F_U0]; a.dudt0 = c->pf[F_DUDT0]; a.u = c->pf[F_U];
  a.segdt = nullptr; a.seg = nullptr;
  k_advance<<<nblocks(c->N, BS), BS, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  c->tree_valid = false;
  return GDF_OK;
}

int integ_wrap(gdf_ctx* c) {
  if (!c->phys.periodic) return GDF_OK;
  WrapArgs w;
  w.N = c->N;
  for (int k = 0; k < 3; k++) {
    w.lo[k] = c->phys.boxmin[k]; w.hi[k] = c->phys.boxmax[k]; w.size[k] = c->phys.boxsize[k];
    w.r[k] = c->pf[F_X + k]; w.r0[k] = c->pf[F_X0 + k];
  }
  k_wrap<<<nblocks(c->N, BS), BS, 0, c->stream>>>(w);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int integ_end_step(gdf_ctx* c, double dtstep) {
  EndArgs e;
  e.N = c->N; e.energy = c->phys.energy_integration; e.dt = dtstep;
  for (int k = 0; k < 3; k++) {
    e.a[k] = c->df[D_AX + k]; e.v[k] = c->pf[F_VX + k]; e.r0[k] = c->pf[F_X0 + k];
    e.v0[k] = c->pf[F_VX0 + k]; e.a0[k] = c->pf[F_AX0 + k]; e.r[k] = c->pf[F_X + k];
  }
  e.dudt = c->df[D_DUDT]; e.u = c->pf[F_U]; e.u0 = c->pf[F_U0]; e.dudt0 = c->pf[F_DUDT0];
  e.segdt = nullptr; e.seg = nullptr;
  k_end_step<<<nblocks(c->N, BS), BS, 0, c->stream>>>(e);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int integ_seg_advance(gdf_ctx* c) {
  AdvArgs a;
  a.N = c->N; a.energy = c->phys.energy_integration; a.dt = 0.0;
  for (int k = 0; k < 3; k++) {
    a.r0[k] = c->pf[F_X0 + k]; a.v0[k] = c->pf[F_VX0 + k]; a.a0[k] = c->pf[F_AX0 + k];
    a.r[k] = c->pf[F_X + k]; a.v[k] = c->pf[F_VX + k];
  }
  a.u0 = c->pf[F_U0]; a.dudt0 = c->pf[F_DUDT0]; a.u = c->pf[F_U];
  a.segdt = c->seg_dt; a.seg = c->pi[I_SEG];
  k_advance<<<nblocks(c->N, BS), BS, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  c->tree_valid = false;
  return GDF_OK;
}

int integ_seg_end_step(gdf_ctx* c) {
  EndArgs e;
  e.N = c->N; e.energy = c->phys.energy_integration; e.dt = 0.0;
  for (int k = 0; k < 3; k++) {
    e.a[k] = c->df[D_AX + k]; e.v[k] = c->pf[F_VX + k]; e.r0[k] = c->pf[F_X0 + k];
    e.v0[k] = c->pf[F_VX0 + k]; e.a0[k] = c->pf[F_AX0 + k]; e.r[k] = c->pf[F_X + k];
  }
  e.dudt = c->df[D_DUDT]; e.u = c->pf[F_U]; e.u0 = c->pf[F_U0]; e.dudt0 = c->pf[F_DUDT0];
  e.segdt = c->seg_dt; e.seg = c->pi[I_SEG];
  k_end_step<<<nblocks(c->N, BS), BS, 0, c->stream>>>(e);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

// SphIntegration::Timestep per particle and Simulation::ComputeGlobalTimestep per segment
int integ_seg_timesteps(gdf_ctx* c, double* out) {
  TsArgs t;
  t.N = c->N; t.hydro = c->phys.hydro_forces; t.energy = c->phys.energy_integration;
  t.nbody = c->phys.nbody;
  t.courant = c->phys.courant_mult; t.emult = c->phys.energy_mult;
  t.accel = c->phys.nbody ? c->phys.nbody_mult : c->phys.accel_mult;
  t.h = c->pf[F_H]; t.sound = c->df[D_SOUND]; t.div_v = c->df[D_DIVV];
  for (int k = 0; k < 3; k++) t.a[k] = c->df[D_AX + k];
  t.u = c->pf[F_U]; t.dudt = c->df[D_DUDT]; t.dt_out = c->df[D_DT]; t.part = c->red;
  const int nb = std::min(nblocks(c->N, BS), RED_BLOCKS);
  k_timestep<<<nb, BS, 0, c->stream>>>(t);
  k_seg_min<<<c->nseg, BS, 0, c->stream>>>(c->d_seg_off, c->df[D_DT], c->red + RED_BLOCKS*ND);
  GDF_HIP(c, hipMemcpyAsync(c->h_red, c->red + RED_BLOCKS*ND, sizeof(double)*c->nseg,
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  for (int sg = 0; sg < c->nseg; sg++) {
    out[sg] = c->h_red[sg];
    if (!(out[sg] > 0.0) || out[sg] > 1.0e20)
      return fail(c, GDF_ERR_TIMESTEP, "Error : Hydro timestep too large");
  }
  return GDF_OK;
}

// per-particle timesteps into D_DT; dtmin == NULL skips reading back their minimum (a block
// step reads it only when it resynchronises the levels)
int integ_timestep(gdf_ctx* c, double* dtmin) {
  if (c->cfg.fixed_dt > 0.0) { if (dtmin) *dtmin = c->cfg.fixed_dt; return GDF_OK; }
  TsArgs t;
  t.N = c->N; t.hydro = c->phys.hydro_forces; t.energy = c->phys.energy_integration;
  t.nbody = c->phys.nbody;
  t.courant = c->phys.courant_mult; t.emult = c->phys.energy_mult;
  t.accel = c->phys.nbody ? c->phys.nbody_mult : c->phys.accel_mult;
  t.h = c->pf[F_H]; t.sound = c->df[D_SOUND]; t.div_v = c->df[D_DIVV];
  for (int k = 0; k < 3; k++) t.a[k] = c->df[D_AX + k];
  t.u = c->pf[F_U]; t.dudt = c->df[D_DUDT]; t.dt_out = c->df[D_DT]; t.part = c->red;
  const int nb = std::min(nblocks(c->N, BS), RED_BLOCKS);
  k_timestep<<<nb, BS, 0, c->stream>>>(t);
  if (!dtmin) return GDF_OK;
  k_min_final<<<1, BS, 0, c->stream>>>(nb, c->red, c->red + RED_BLOCKS*ND);
  GDF_HIP(c, hipMemcpyAsync(c->h_red, c->red + RED_BLOCKS*ND, sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  *dtmin = c->h_red[0];
  GDF_TRY(stars_timestep(c, dtmin));     // Simulation::ComputeGlobalTimestep's star minimum
  if (!(*dtmin > 0.0) || *dtmin > 1.0e20)
    return fail(c, GDF_ERR_TIMESTEP, c->phys.nbody ? "Error : N-body timestep invalid"
                                                   : "Error : Hydro timestep too large");
  return GDF_OK;
}

int integ_max_h(gdf_ctx* c, double* hmax) {
  const int nb = std::min(nblocks(c->N, BS), RED_BLOCKS);
  k_max_h<<<nb, BS, 0, c->stream>>>(c->N, c->pf[F_H], c->red);
  GDF_HIP(c, hipMemcpyAsync(c->h_red, c->red, sizeof(double)*nb, hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  double best = 0.0;
  for (int b = 0; b < nb; b++) best = std::max(best, c->h_red[b]);
  *hmax = best;
  return GDF_OK;
}

int integ_diagnostics(gdf_ctx* c, double* out) {
  DiagArgs d;
  d.N = c->N; d.stride = 1;
  for (int k = 0; k < 3; k++) { d.r[k] = c->pf[F_X + k]; d.v[k] = c->pf[F_VX + k]; }
  d.m = c->pf[F_M]; d.u = c->pf[F_U]; d.gpot = c->df[D_GPOT]; d.part = c->red;
  const int nb = std::min(nblocks(c->N, BS), RED_BLOCKS);
  k_diag<<<nb, BS, 0, c->stream>>>(d);
  GDF_HIP(c, hipMemcpyAsync(c->h_red, c->red, sizeof(double)*ND*nb, hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  double acc[ND] = {};
  for (int b = 0; b < nb; b++) for (int k = 0; k < ND; k++) acc[k] += c->h_red[b*ND + k];
  GDF_TRY(stars_diagnostics(c, acc));    // the star loop after the gas (SimAnalysis.hpp:116-143)
  const double ketot = 0.5*acc[1], gpetot = 0.5*acc[3];
  double Etot = ketot;
  if (c->phys.hydro_forces) Etot += acc[2];
  if (c->phys.self_gravity || c->Nstar > 0) Etot += gpetot;
  out[0] = Etot; out[1] = ketot; out[2] = acc[2]; out[3] = gpetot;
  for (int k = 0; k < 3; k++) { out[4 + k] = acc[4 + k]; out[7 + k] = acc[7 + k]; }
  out[10] = acc[0];
  return GDF_OK;
}

void fill_int(gdf_ctx* c, int* p, int n, int v) {
  if (n > 0) k_fill_i<<<nblocks(n, BS), BS, 0, c->stream>>>(n, p, v);
}

// Sph::ZeroAccelerations (Sph.cpp:126-140): every particle with global timesteps, the active
// ones (and their neighbour levels) with block timesteps, whose thermal state is refreshed for
// every particle as SphSimulation::MainLoop does (SphSimulation.cpp:667-676, nradstep = 1)
int integ_zero_accelerations(gdf_ctx* c) {
  if (c->nlevels == 1) return integ_zero_forces(c);
  GDF_TRY(block_zero_active(c));
  return block_eos_all(c);
}

int integ_zero_forces(gdf_ctx* c) {
  const int n = c->N;
  const int fields[] = {D_AX, D_AY, D_AZ, D_ATX, D_ATY, D_ATZ, D_DUDT, D_DIVV, D_GPOT};
  for (int f : fields) k_zero_d<<<nblocks(n, BS), BS, 0, c->stream>>>(n, c->df[f]);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int integ_add_atree(gdf_ctx* c) {
  k_add_atree<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, c->df[D_AX], c->df[D_AY],
                                                        c->df[D_AZ], c->df[D_ATX], c->df[D_ATY],
                                                        c->df[D_ATZ],
                                                        c->nlevels > 1 ? c->pi[I_ACT] : nullptr);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int integ_pack_posm(gdf_ctx* c) {
  k_pack_posm<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, c->pf[F_X], c->pf[F_Y],
                                                        c->pf[F_Z], c->pf[F_M], c->posm);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

// ---------------------------------------------------------------------------------------------
// Block timesteps: SphLeapfrogKDK::AdvanceParticles / CheckTimesteps / EndTimestep with per-
// particle steps (SphLeapfrogKDK.cpp:76-329) and Simulation::ComputeBlockTimesteps
// (Simulation.cpp:1764-2197) for hydro particles, KDK (integration_step = 1), no sinks,
// sph_single_timestep = 0.  Integer state per particle: level, levelneib, nstep, nlast, and the
// active / end_timestep flags; the integer time n, level_max, level_step, nresync and dt_max
// are host scalars, updated from two small reductions per step.
// ---------------------------------------------------------------------------------------------
namespace {

// (int) pow(2, e) as the reference computes step lengths: 0 for negative e
__device__ __host__ __forceinline__ int ipow2(int e) { return e >= 0 ? (1 << e) : 0; }

// ComputeTimestepLevel (InlineFuncs.h:550-557)
__device__ __forceinline__ int timestep_level(double dt, double dt_max) {
  const int level = (int) (1.44269504088896*log(dt_max/dt)) + 1;
  return level > 0 ? level : 0;
}

struct BlkArgs {
  int N, n, level_step, level_max, ldm, energy;
  double t, timestep, dt_max;
  double* r[3]; double* v[3]; const double* r0[3]; const double* v0[3]; const double* a0[3];
  double* u; const double* u0; const double* dudt0;
  double* tlast; double* pdt; double* dtnext;
  int* level; int* levelneib; int* nstep; int* nlast; int* act; int* end;
  const double* dtw;
  int* counter; long long* finished;
};

__global__ void k_block_init(BlkArgs b) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= b.N) return;
  b.level[i] = 0; b.levelneib[i] = 0; b.nstep[i] = 0; b.nlast[i] = 0;
  b.act[i] = 1; b.end[i] = 0;
  b.tlast[i] = b.t; b.pdt[i] = 0.0; b.dtnext[i] = 0.0;
}

// drift every particle from the start of its own step; active = at the end of its step
__global__ void k_block_advance(BlkArgs b) {
#pragma clang fp contract(off)
  long long na = 0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < b.N; i += gridDim.x*BS) {
    const double dt = b.t - b.tlast[i];
    for (int k = 0; k < 3; k++) b.r[k][i] = b.r0[k][i] + b.v0[k][i]*dt + 0.5*b.a0[k][i]*dt*dt;
    for (int k = 0; k < 3; k++) b.v[k][i] = b.v0[k][i] + b.a0[k][i]*dt;
    if (b.energy) b.u[i] = b.u0[i] + b.dudt0[i]*dt;
    const int a = (b.n - b.nlast[i] == b.nstep[i]) ? 1 : 0;
    b.act[i] = a;
    na += a;
  }
  block_add_ll(reinterpret_cast<long long*>(b.counter), na);   // active particles of the step
}

// Sph::ZeroAccelerations (Sph.cpp:126-140) for the active particles
struct ZeroArgs { int N; const int* act; int* levelneib; double* f[9]; };
__global__ void k_zero_active(ZeroArgs z) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= z.N || !z.act[i]) return;
  z.levelneib[i] = 0;
  #pragma unroll
  for (int k = 0; k < 9; k++) z.f[k][i] = 0.0;
}

// ComputeThermalProperties for every particle (SphSimulation.cpp:667-676 with nradstep = 1)
__global__ void k_eos_all(int N, PhysParams p, const double* __restrict__ rho,
                          double* __restrict__ u, double* __restrict__ sound,
                          double* __restrict__ pressure) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  double uu = u[i], snd, press;
  eos_state(p, rho[i], uu, snd, press);
  u[i] = uu; sound[i] = snd; pressure[i] = press;
}

// SphLeapfrogKDK::CheckTimesteps: end a step early where a neighbour's level is too far above;
// every particle's active flag is rewritten (set only for these), counter += their number
__global__ void k_check_timesteps(BlkArgs b) {
  long long newly = 0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < b.N; i += gridDim.x*BS) {
    int a = 0;
    const int dn = b.n - b.nlast[i];
    const int level = b.level[i], ln = b.levelneib[i];
    if (dn != b.nstep[i] && ln - level > b.ldm) {
      const int level_new = ln - b.ldm;
      const int nnewstep = ipow2(b.level_step - level_new);
      if (nnewstep > 0 && dn%nnewstep == 0) {
        if (dn > 0) b.nstep[i] = dn;
        b.level[i] = level_new;
        a = 1;
      }
    }
    b.act[i] = a;
    newly += a;
  }
  block_add_ll(reinterpret_cast<long long*>(b.counter), newly);
}

// resynchronisation branch: every particle gets a fresh level from its timestep
__global__ void k_block_resync(BlkArgs b) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= b.N) return;
  const int level = min(timestep_level(b.dtw[i], b.dt_max), b.level_max);
  b.level[i] = level;
  b.levelneib[i] = level;
  const int nstep = ipow2(b.level_step - level);
  b.nstep[i] = nstep;
  b.nlast[i] = b.n;
  b.dtnext[i] = nstep*b.timestep;
  b.end[i] = 1;
}

// particles at the end of their step choose their next level; counter = max level of all
__global__ void k_block_update(BlkArgs b) {
  int lv = 0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < b.N; i += gridDim.x*BS) {
    const int n = b.n;
    int level = b.level[i];
    const int nstep = b.nstep[i], nlast = b.nlast[i];
    if (n - nlast == nstep) {
      int lnew = timestep_level(b.dtw[i], b.dt_max);
      lnew = max(lnew, b.levelneib[i] - b.ldm);
      if (nstep != ipow2(b.level_step - level)) {
        // step ended early by CheckTimesteps
        level = max(level, lnew);
        b.levelneib[i] = level;
      } else {
        if (lnew < level && level > 1 && n%(2*nstep) == 0) level = level - 1;
        else if (lnew > level) level = lnew;
        b.levelneib[i] = lnew;
      }
      b.level[i] = level;
      b.nlast[i] = n;
      const int ns = ipow2(b.level_step - level);
      b.nstep[i] = ns;
      b.dtnext[i] = ns*b.timestep;
      b.end[i] = 1;
    }
    lv = max(lv, level);
  }
  // one atomic per workgroup
  __shared__ int part[BS/64];
  lv = wave_max_int(lv);
  if (__lane_id() == 0) part[threadIdx.x/64] = lv;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < BS/64; w++) lv = max(lv, part[w]);
    atomicMax(b.counter, lv);
  }
}

// integer time rescaled when level_max changes: mul = 1 multiplies nstep and nlast by f, else
// divides them
__global__ void k_block_rescale(BlkArgs b, int f, int mul) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= b.N) return;
  if (mul) { b.nstep[i] *= f; b.nlast[i] *= f; }
  else { b.nlast[i] /= f; b.nstep[i] /= f; }
}

__global__ void k_block_fix_nstep(BlkArgs b) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= b.N) return;
  if (b.nlast[i] == b.n) b.nstep[i] = ipow2(b.level_step - b.level[i]);
}

// SphLeapfrogKDK::EndTimestep for the particles flagged end_timestep
struct BlkEndArgs {
  int N, n, energy;
  double t;
  const double* a[3]; double* v[3]; double* r0[3]; double* v0[3]; double* a0[3];
  const double* r[3];
  const double* dudt; double* u; double* u0; double* dudt0;
  double* tlast; double* pdt; double* dtnext;
  int* nlast; int* act; int* end;
  long long* finished;
};

__global__ void k_block_end_step(BlkEndArgs e) {
#pragma clang fp contract(off)
  long long fin = 0;
  for (int i = blockIdx.x*BS + threadIdx.x; i < e.N; i += gridDim.x*BS) {
    if (!e.end[i]) continue;
    fin++;
    const double dt = e.pdt[i];
    for (int k = 0; k < 3; k++) {
      const double a = e.a[k][i], a0 = e.a0[k][i];
      const double v = e.v[k][i] + 0.5*dt*(a - a0);
      e.v[k][i] = v;
      e.r0[k][i] = e.r[k][i];
      e.v0[k][i] = v;
      e.a0[k][i] = a;
    }
    if (e.energy) {
      const double dudt = e.dudt[i], dudt0 = e.dudt0[i];
      double u = e.u[i] + 0.5*(dudt - dudt0)*dt;
      if (u <= 0.0) u = e.u0[i] + dudt0*dt;
      e.u[i] = u;
      e.u0[i] = u;
      e.dudt0[i] = dudt;
    }
    e.nlast[i] = e.n;
    e.tlast[i] = e.t;
    e.pdt[i] = e.dtnext[i];
    e.dtnext[i] = 0.0;
    e.act[i] = 0;
    e.end[i] = 0;
  }
  block_add_ll(e.finished, fin);
}

// grid of the passes that count into one counter: at most 1024 workgroups, one atomic each
int count_grid(int n) { return std::min(nblocks(n, BS), 1024); }

BlkArgs blk_args(gdf_ctx* c) {
  BlkArgs b;
  b.N = c->N; b.n = c->n_int; b.level_step = c->level_step; b.level_max = c->level_max;
  b.ldm = c->level_diff_max; b.energy = c->phys.energy_integration;
  b.t = c->t; b.timestep = c->timestep; b.dt_max = c->dt_max;
  for (int k = 0; k < 3; k++) {
    b.r[k] = c->pf[F_X + k]; b.v[k] = c->pf[F_VX + k];
    b.r0[k] = c->pf[F_X0 + k]; b.v0[k] = c->pf[F_VX0 + k]; b.a0[k] = c->pf[F_AX0 + k];
  }
  b.u = c->pf[F_U]; b.u0 = c->pf[F_U0]; b.dudt0 = c->pf[F_DUDT0];
  b.tlast = c->pf[F_TLAST]; b.pdt = c->pf[F_PDT]; b.dtnext = c->pf[F_DTNEXT];
  b.level = c->pi[I_LEVEL]; b.levelneib = c->pi[I_LEVELNEIB]; b.nstep = c->pi[I_NSTEP];
  b.nlast = c->pi[I_NLAST]; b.act = c->pi[I_ACT]; b.end = c->pi[I_END];
  b.dtw = c->df[D_DT];
  b.counter = c->counters + 12;
  b.finished = c->icount + 6;
  return b;
}

int read_counter(gdf_ctx* c, int* out) {
  GDF_HIP(c, hipMemcpyAsync(c->h_counters + 12, c->counters + 12, sizeof(int)*2,
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  *out = c->h_counters[12];
  return GDF_OK;
}

__global__ void k_zero_counter(int* __restrict__ p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

int zero_counter(gdf_ctx* c) {
  // a kernel rather than hipMemsetAsync: the runtime's memset path costs more host time per call
  k_zero_counter<<<1, 64, 0, c->stream>>>(c->counters + 12);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

}  // namespace

__global__ void k_zero_icount(long long* __restrict__ p) {
  if (threadIdx.x < 8) p[threadIdx.x] = 0;
}

int zero_icount(gdf_ctx* c) {
  k_zero_icount<<<1, 64, 0, c->stream>>>(c->icount);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int block_init(gdf_ctx* c) {
  c->n_int = 0; c->nresync = 0; c->level_max = 0; c->level_step = 0; c->dt_max = 0.0;
  c->n_active = c->N;                  // the setup passes compute every particle
  BlkArgs b = blk_args(c);
  k_block_init<<<nblocks(c->N, BS), BS, 0, c->stream>>>(b);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int block_advance(gdf_ctx* c) {
  GDF_TRY(zero_counter(c));
  BlkArgs b = blk_args(c);
  k_block_advance<<<count_grid(c->N), BS, 0, c->stream>>>(b);
  GDF_HIP(c, hipGetLastError());
  // read with the tree build's synchronisation
  GDF_HIP(c, hipMemcpyAsync(c->h_counters + 12, c->counters + 12, sizeof(int)*2,
                            hipMemcpyDeviceToHost, c->stream));
  c->tree_valid = false;
  return GDF_OK;
}

int block_zero_active(gdf_ctx* c) {
  ZeroArgs z;
  z.N = c->N; z.act = c->pi[I_ACT]; z.levelneib = c->pi[I_LEVELNEIB];
  const int fields[9] = {D_AX, D_AY, D_AZ, D_ATX, D_ATY, D_ATZ, D_DUDT, D_DIVV, D_GPOT};
  for (int k = 0; k < 9; k++) z.f[k] = c->df[fields[k]];
  k_zero_active<<<nblocks(c->N, BS), BS, 0, c->stream>>>(z);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int block_eos_all(gdf_ctx* c) {
  k_eos_all<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, c->phys, c->df[D_RHO], c->pf[F_U],
                                                      c->df[D_SOUND], c->df[D_PRESSURE]);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int block_check_timesteps(gdf_ctx* c, int* activecount) {
  GDF_TRY(zero_counter(c));
  BlkArgs b = blk_args(c);
  k_check_timesteps<<<count_grid(c->N), BS, 0, c->stream>>>(b);
  GDF_HIP(c, hipGetLastError());
  return read_counter(c, activecount);
}

int block_timesteps(gdf_ctx* c) {
  double dmin = 0.0;
  // per-particle criteria into D_DT; their minimum only where the levels resynchronise
  const bool resync = c->n_int == c->nresync;
  GDF_TRY(integ_timestep(c, resync ? &dmin : nullptr));
  if (resync) {
    c->n_int = 0;
    c->timestep = dmin;
    c->level_max = c->nlevels - 1;
    c->level_step = c->level_max;
    c->dt_max = c->timestep*std::pow(2.0, c->level_max);
    BlkArgs b = blk_args(c);
    k_block_resync<<<nblocks(c->N, BS), BS, 0, c->stream>>>(b);
    GDF_HIP(c, hipGetLastError());
    c->nresync = ipow2(c->level_step);
    c->timestep = c->dt_max/(double) c->nresync;
    return GDF_OK;
  }
  const int level_max_old = c->level_max;
  GDF_TRY(zero_counter(c));
  BlkArgs b = blk_args(c);
  k_block_update<<<count_grid(c->N), BS, 0, c->stream>>>(b);
  GDF_HIP(c, hipGetLastError());
  int level_max = 0;
  GDF_TRY(read_counter(c, &level_max));
  const int istep = ipow2(c->level_step - level_max_old + 1);
  if (level_max > level_max_old) {
    const int nfactor = ipow2(level_max - level_max_old);
    c->n_int *= nfactor;
    k_block_rescale<<<nblocks(c->N, BS), BS, 0, c->stream>>>(b, nfactor, 1);
  } else if (level_max <= level_max_old - 1 && level_max_old > 1 && c->n_int%istep == 0) {
    level_max = level_max_old - 1;
    const int nfactor = ipow2(level_max_old - level_max);
    c->n_int /= nfactor;
    k_block_rescale<<<nblocks(c->N, BS), BS, 0, c->stream>>>(b, nfactor, 0);
  } else {
    level_max = level_max_old;
  }
  GDF_HIP(c, hipGetLastError());
  c->level_max = level_max;
  c->level_step = level_max;
  c->nresync = ipow2(c->level_step);
  c->timestep = c->dt_max/(double) c->nresync;
  b = blk_args(c);
  k_block_fix_nstep<<<nblocks(c->N, BS), BS, 0, c->stream>>>(b);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int block_end_step(gdf_ctx* c) {
  BlkEndArgs e;
  e.N = c->N; e.n = c->n_int; e.energy = c->phys.energy_integration; e.t = c->t;
  for (int k = 0; k < 3; k++) {
    e.a[k] = c->df[D_AX + k]; e.v[k] = c->pf[F_VX + k]; e.r0[k] = c->pf[F_X0 + k];
    e.v0[k] = c->pf[F_VX0 + k]; e.a0[k] = c->pf[F_AX0 + k]; e.r[k] = c->pf[F_X + k];
  }
  e.dudt = c->df[D_DUDT]; e.u = c->pf[F_U]; e.u0 = c->pf[F_U0]; e.dudt0 = c->pf[F_DUDT0];
  e.tlast = c->pf[F_TLAST]; e.pdt = c->pf[F_PDT]; e.dtnext = c->pf[F_DTNEXT];
  e.nlast = c->pi[I_NLAST]; e.act = c->pi[I_ACT]; e.end = c->pi[I_END];
  e.finished = c->icount + 6;
  k_block_end_step<<<count_grid(c->N), BS, 0, c->stream>>>(e);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

}  // namespace gdf

// ---- diagnostics with no device allocation ----------------------------------------------------
// The block partials of k_diag live in a file-scope device array instead of a context's
// reduction buffer: RED_BLOCKS x ND doubles (88 KiB), held by every GPU the library runs on
// until the process ends.  Host code never passes the array's name to a kernel (on the host it
// names a stand-in, not device memory): the address comes from hipGetSymbolAddress, looked up
// once per device with that device current, and kept per device.  Every call writes the
// partials it reads (blocks 0..nb-1), so nothing needs clearing between calls; calls that
// overlap on one GPU share the array and race.
namespace gdf {
namespace {

__device__ __align__(256) double scratch[RED_BLOCKS*ND];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};
thread_local std::string g_noalloc_err;

int noalloc_fail(int code, const char* msg) {
  g_noalloc_err = msg;
  return code;
}

// device address of this device's scratch array (looked up on the first call on the device)
int scratch_on_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return noalloc_fail(GDF_ERR_HIP, "gdf_calculate_diagnostics_noalloc: no current device");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return noalloc_fail(GDF_ERR_HIP, "gdf_calculate_diagnostics_noalloc: "
                                       "hipGetSymbolAddress of the scratch array failed");
    g_scratch_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[dev];
  return GDF_OK;
}

// a pointer a kernel on this device may read or write: device, pinned host or managed memory
bool device_accessible(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void) hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeHost ||
         a.type == hipMemoryTypeManaged || a.type == hipMemoryTypeUnified;
}

}  // namespace

const char* noalloc_last_error() {
  return g_noalloc_err.empty() ? nullptr : g_noalloc_err.c_str();
}

}  // namespace gdf

extern "C" int gdf_calculate_diagnostics_noalloc(int N, const double* r, const double* v,
                                                 const double* m, const double* u,
                                                 const double* gpot, int hydro_forces,
                                                 int self_gravity, double* out, void* stream) {
  using namespace gdf;
  if (N < 1 || (long long) N > kNoallocMaxN)
    return noalloc_fail(GDF_ERR_INVALID, "gdf_calculate_diagnostics_noalloc: N must be in "
                                         "[1, 2^30]");
  const void* ptrs[] = {r, v, m, u, gpot, out};
  for (const void* p : ptrs)
    if (!p || !device_accessible(p))
      return noalloc_fail(GDF_ERR_INVALID, "gdf_calculate_diagnostics_noalloc: every array "
                                           "must be device-accessible memory");
  double* part = nullptr;
  GDF_TRY(scratch_on_current_device(&part));
  hipStream_t s = static_cast<hipStream_t>(stream);
  DiagArgs d;
  d.N = N; d.stride = 3;
  for (int k = 0; k < 3; k++) { d.r[k] = r + k; d.v[k] = v + k; }
  d.m = m; d.u = u; d.gpot = gpot; d.part = part;
  const int nb = std::min(nblocks(N, BS), RED_BLOCKS);
  k_diag<<<nb, BS, 0, s>>>(d);
  k_diag_final<<<1, 64, 0, s>>>(nb, part, hydro_forces != 0, self_gravity != 0, out);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return noalloc_fail(GDF_ERR_HIP, "gdf_calculate_diagnostics_noalloc: kernel launch failed");
  g_noalloc_err.clear();
  return GDF_OK;
}
