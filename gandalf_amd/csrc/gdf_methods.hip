// Per-method entry points of the reference's integrator and neighbour-search interfaces, for a
// caller that drives the step itself (a reference-side SphSimulation::MainLoop, including its
// block timesteps), rather than through gdf_main_loop:
//
//   TimeIntegration<ndim> (src/Headers/Integration.h:48-66), SphLeapfrogKDK
//   (src/Hydrodynamics/SphLeapfrogKDK.cpp): AdvanceParticles(n, t, timestep),
//   CorrectionTerms, SetActiveParticles, CheckTimesteps, EndTimestep(n, t, timestep);
//   Simulation::ComputeBlockTimesteps (src/Common/Simulation.cpp:1764-2197);
//   NeighbourSearch<ndim>::GetGatherNeighbourList (src/Headers/NeighbourSearch.h:83,
//   HydroTree::GetGatherNeighbourList, src/Tree/HydroTree.cpp:451-471).
//
// The integer time n, the time t and the base timestep are the caller's, as in the reference's
// signatures: each entry point sets the context's block clock from its arguments.  The kernels
// are streaming passes over the particle arrays (HBM-bound).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;

// SphLeapfrogKDK::SetActiveParticles (SphLeapfrogKDK.cpp:192-211)
__global__ void k_set_active(int N, int n, const int* __restrict__ nlast,
                             const int* __restrict__ nstep, const int* __restrict__ end,
                             int* __restrict__ act) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  act[i] = (end[i] || n - nlast[i] == nstep[i]) ? 1 : 0;
}

struct CorrArgs {
  int N, n, energy, global;
  double timestep;
  const int* nlast; const int* nstep; const double* pdt;
  const double* a[3]; const double* a0[3]; double* v[3];
  const double* dudt; double* dudt0; double* u; double* u0;
};

// SphLeapfrogKDK::CorrectionTerms (SphLeapfrogKDK.cpp:138-183): particles at the end of their
// step get the second-order velocity (and internal energy) correction.  With global timesteps
// every particle ends its step and its step length is the base timestep.
__global__ void k_correction_terms(CorrArgs c) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= c.N) return;
  if (!c.global && c.n - c.nlast[i] != c.nstep[i]) return;
  const double dt = c.global ? c.timestep : c.pdt[i];
  for (int k = 0; k < 3; k++) c.v[k][i] += 0.5*dt*(c.a[k][i] - c.a0[k][i]);
  if (c.energy) {
    const double dudt = c.dudt[i], dudt0 = c.dudt0[i];
    double u = c.u[i] + 0.5*(dudt - dudt0)*dt;
    if (u <= 0.0) u = c.u0[i] + dudt0*dt;
    c.u[i] = u;
    c.u0[i] = u;
    c.dudt0[i] = dudt;
  }
}

// Tree::ComputeGatherNeighbourList, point variant (Tree.cpp:208-280): the particles with
// |r_i - rp|^2 < rsearch^2 (strict), by one pass over the positions; periodic boxes take the
// nearest image, as the reference's ghost tree supplies the periodic copies.  Matches append
// their import index through a counter (order is fixed afterwards on the host).
__global__ void k_gather_point(int N, const double* __restrict__ x, const double* __restrict__ y,
                               const double* __restrict__ z, double px, double py, double pz,
                               double rs2, PhysParams p, const int* __restrict__ iorig,
                               int* __restrict__ out, int cap, int* __restrict__ count) {
#pragma clang fp contract(off)
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  double dx = x[i] - px, dy = y[i] - py, dz = z[i] - pz;
  nearest_image<double>(p, dx, dy, dz);
  if (!(dot3_rn(dx, dy, dz) < rs2)) return;
  const int k = atomicAdd(count, 1);
  if (k < cap) out[k] = iorig[i];
}

}  // namespace
}  // namespace gdf

using namespace gdf;

extern "C" {

int gdf_set_active_particles(gdf_ctx* c, int n) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (c->nlevels == 1) {
    // one global timestep: every particle ends its step at every integer time
    fill_int(c, c->pi[I_ACT], c->N, 1);
  } else {
    c->n_int = n;
    k_set_active<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, n, c->pi[I_NLAST],
                                                          c->pi[I_NSTEP], c->pi[I_END],
                                                          c->pi[I_ACT]);
  }
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_correction_terms(gdf_ctx* c, int n, double t, double timestep) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  (void) t;
  CorrArgs a;
  a.N = c->N; a.n = n; a.energy = c->phys.energy_integration; a.global = c->nlevels == 1;
  a.timestep = timestep;
  a.nlast = c->pi[I_NLAST]; a.nstep = c->pi[I_NSTEP]; a.pdt = c->pf[F_PDT];
  for (int k = 0; k < 3; k++) {
    a.a[k] = c->df[D_AX + k]; a.a0[k] = c->pf[F_AX0 + k]; a.v[k] = c->pf[F_VX + k];
  }
  a.dudt = c->df[D_DUDT]; a.dudt0 = c->pf[F_DUDT0]; a.u = c->pf[F_U]; a.u0 = c->pf[F_U0];
  k_correction_terms<<<nblocks(c->N, BS), BS, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_check_timesteps(gdf_ctx* c, int level_diff_max, int level_step, int n, double timestep,
                        int* activecount) {
  if (!c || c->N <= 0 || !activecount || level_diff_max < 1) return GDF_ERR_INVALID;
  *activecount = 0;
  if (c->nlevels == 1) return GDF_OK;        // the reference checks only with Nlevels > 1
  c->level_diff_max = level_diff_max;
  c->level_step = level_step;
  c->n_int = n;
  c->timestep = timestep;
  return block_check_timesteps(c, activecount);
}

int gdf_compute_block_timesteps(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (c->nlevels == 1) return fail(c, GDF_ERR_INVALID, "block timesteps need Nlevels > 1");
  GDF_TRY(block_timesteps(c));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_advance_particles_block(gdf_ctx* c, int n, double t, double timestep) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (c->nlevels == 1) {
    GDF_TRY(integ_advance(c, t - c->tlast));
    fill_int(c, c->pi[I_ACT], c->N, 1);
  } else {
    c->n_int = n;
    c->t = t;
    c->timestep = timestep;
    GDF_TRY(block_advance(c));
  }
  c->t = t;
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_end_timestep_block(gdf_ctx* c, int n, double t, double timestep) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (c->nlevels == 1) {
    GDF_TRY(integ_end_step(c, timestep));
  } else {
    c->n_int = n;
    c->t = t;
    GDF_TRY(block_end_step(c));
  }
  c->tlast = t;
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_get_gather_neighbour_list(gdf_ctx* c, const double* rp, double rsearch, int nmax,
                                  int* list) {
  if (!c || c->N <= 0 || !rp || !list || nmax < 0 || !(rsearch >= 0.0))
    return GDF_ERR_INVALID;
  // the match buffer is the context's export scratch (3N doubles: room for 6N ids)
  int* buf = reinterpret_cast<int*>(c->tmpd);
  int* count = c->counters + 14;
  fill_int(c, count, 1, 0);
  k_gather_point<<<nblocks(c->N, BS), BS, 0, c->stream>>>(
      c->N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z], rp[0], rp[1], rp[2], rsearch*rsearch,
      c->phys, c->iorig, buf, c->N, count);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipMemcpyAsync(c->h_counters + 14, count, sizeof(int), hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  const int nfound = c->h_counters[14];
  if (nfound > nmax) return -1;     // the reference's overflow code: the caller grows its buffer
  std::vector<int> ids(nfound);
  if (nfound > 0)
    GDF_HIP(c, hipMemcpy(ids.data(), buf, sizeof(int)*nfound, hipMemcpyDeviceToHost));
  std::sort(ids.begin(), ids.end());
  std::copy(ids.begin(), ids.end(), list);
  return nfound;
}

}  // extern "C"
