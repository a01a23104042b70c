// Host driver of the particle hot path; see gradhsph_simulation_gpu.h.
#include "gradhsph_simulation_gpu.h"

#include <cmath>
#include <vector>

namespace gandalf_gpu {

namespace {

// Sph::InitialSmoothingLengthGuess, 3D branch (src/Hydrodynamics/Sph.cpp:76-113): a uniform
// medium filling the particles' bounding box; powf as in the reference
double initial_h_guess(const std::vector<double>& r, double kernrange, double h_fac) {
  const size_t N = r.size()/3;
  double rmin[3] = {9.9e20, 9.9e20, 9.9e20}, rmax[3] = {-9.9e20, -9.9e20, -9.9e20};
  for (size_t i = 0; i < N; i++)
    for (int k = 0; k < 3; k++) {
      if (r[3*i + k] < rmin[k]) rmin[k] = r[3*i + k];
      if (r[3*i + k] > rmax[k]) rmax[k] = r[3*i + k];
    }
  const double pi = 3.14159265358979;
  const int Ngather = (int) (4.0*pi*std::pow(kernrange*h_fac, 3)/3.0);
  const double volume = (rmax[0] - rmin[0])*(rmax[1] - rmin[1])*(rmax[2] - rmin[2]);
  return powf((3.0*volume*(double) Ngather)/(32.0*pi*(double) N), 0.33333333333333333333333);
}

void record(gdf_ctx* c, int e) {
  if (c->timing && c->mark_passes) hipEventRecord(c->ev[e], c->stream);
}

}  // namespace

template <int ndim>
int GradhSphTreeGpu<ndim>::BuildTree(bool rebuild_tree, int n, int ntreebuildstep,
                                     int ntreestockstep, double) {
  const int rc = gdf_build_tree_step(ctx, rebuild_tree ? 1 : 0, n, ntreebuildstep,
                                     ntreestockstep);
  return rc < 0 ? rc : GDF_OK;
}

template <int ndim>
int GradhSphTreeGpu<ndim>::UpdateAllSphProperties(bool nopbc) {
  return gdf::sph_density(ctx, nopbc ? 1 : 0);
}

static int zero_accelerations(gdf_ctx* c) { return gdf::integ_zero_accelerations(c); }

template <int ndim>
int GradhSphTreeGpu<ndim>::UpdateAllSphHydroForces() {
  GDF_TRY(zero_accelerations(ctx));
  GDF_TRY(gdf::sph_build_sg(ctx));
  GDF_TRY(gdf::sph_hydro(ctx, 0));
  // the stars' pull on the active gas (GradhSphTree.cpp:385-391)
  return gdf::stars_on_gas(ctx);
}

template <int ndim>
int GradhSphTreeGpu<ndim>::UpdateAllSphForces() {
  GDF_TRY(zero_accelerations(ctx));
  if (ctx->phys.hydro_forces) {
    // the gravity walk needs only h and zeta of the density pass: it runs on the second stream
    // while the scatter-gather walk and the hydro pass run on the first
    GDF_TRY(gdf::gravity_launch(ctx));
    GDF_TRY(gdf::sph_build_sg(ctx));
    GDF_TRY(gdf::sph_hydro(ctx, 1));
    record(ctx, 3);
    GDF_TRY(gdf::gravity_join(ctx));
    return gdf::stars_on_gas(ctx);      // GradhSphTree.cpp:600-607
  }
  record(ctx, 3);
  GDF_TRY(gdf::gravity_walk(ctx));
  GDF_TRY(gdf::integ_add_atree(ctx));
  return gdf::stars_on_gas(ctx);
}

template <int ndim>
int GradhSphTreeGpu<ndim>::GetMaximumSmoothingLength(double* hmax) {
  return gdf::integ_max_h(ctx, hmax);
}

template <int ndim>
int SphLeapfrogKDKGpu<ndim>::AdvanceParticles(int, double t, double) {
  GDF_TRY(gdf::integ_advance(ctx, t - ctx->tlast));
  // NbodyLeapfrogKDK::AdvanceParticles beside the gas's (SphSimulation.cpp:594-595)
  return gdf::stars_advance(ctx, t);
}

template <int ndim>
int SphLeapfrogKDKGpu<ndim>::CheckBoundaries() {
  return gdf::integ_wrap(ctx);
}

template <int ndim>
int SphLeapfrogKDKGpu<ndim>::EndTimestep(int, double t, double timestep) {
  GDF_TRY(gdf::integ_end_step(ctx, timestep));
  // the stars' end-step terms after the gas's (SphSimulation.cpp:874-881)
  GDF_TRY(gdf::stars_end_step(ctx, t));
  ctx->tlast = t;
  return GDF_OK;
}

template <int ndim>
int GradhSphSimulationGpu<ndim>::ComputeGlobalTimestep(double* dt) {
  return gdf::integ_timestep(ctx, dt);
}

// The N-body iteration of the step (SphSimulation.cpp:775-807, Npec = 1): the stars' a, adot
// and gpot cleared, the gas's pull (self-gravitating runs), the star-star sums, and the
// velocity correction for the step that ends at t
static int star_forces(gdf_ctx* c, bool correct) {
  if (c->Nstar <= 0) return GDF_OK;
  GDF_TRY(gdf::stars_gas_on_stars(c, c->phys.self_gravity && c->N > 0, 0));
  GDF_TRY(gdf::nbody_direct_stars(c, c->cfg.nbody_softening));
  return correct ? gdf::stars_correction(c, c->t) : GDF_OK;
}

template <int ndim>
int GradhSphSimulationGpu<ndim>::ComputeForces() {
  if (ctx->phys.self_gravity) return sphneib.UpdateAllSphForces();
  if (ctx->phys.hydro_forces) {
    GDF_TRY(sphneib.UpdateAllSphHydroForces());
    record(ctx, 3);
    return GDF_OK;
  }
  return gdf::fail(ctx, GDF_ERR_INVALID, "Error: No forces included in simulation");
}

// The setup starts from the time and step count the particles came with: zero after
// gdf_set_particles, the file's after a snapshot read, which the reference keeps through its
// setup (SimulationIO.hpp:1363-1382); the file's particle ids survive only a restart
// (SphSimulation.cpp:215-217)
static void initial_state(gdf_ctx* c, double* t0, int64_t* nsteps0) {
  *t0 = c->t;
  *nsteps0 = c->nsteps;
  if (!c->cfg.restart) c->snap_porig.clear();
}

template <int ndim>
int GradhSphSimulationGpu<ndim>::PostInitialConditionsSetup() {
  gdf_ctx* c = ctx;
  double t0;
  int64_t nsteps0;
  initial_state(c, &t0, &nsteps0);
  // initial smoothing lengths (Sph::InitialSmoothingLengthGuess), then the three h-rho passes
  // of the reference setup: before periodic ghosts exist, after them, and in the force loop
  std::vector<double> r(3*(size_t) c->N);
  GDF_TRY(gdf_get_field(c, "r", r.data()));
  const double hg = initial_h_guess(r, c->phys.kern.range, c->phys.h_fac);
  std::vector<double> h((size_t) c->N, hg);
  GDF_TRY(gdf_set_field(c, "h", h.data()));
  c->t = t0;
  // every particle active, level 0, tlast = t (SphSimulation.cpp:225-330)
  if (c->nlevels > 1) GDF_TRY(gdf::block_init(c));
  GDF_TRY(sphneib.BuildTree(true, 0, 1, 1, 0.0));
  GDF_TRY(sphneib.UpdateAllSphProperties(c->phys.periodic != 0));
  GDF_TRY(sphneib.BuildTree(true, 0, 1, 1, 0.0));
  GDF_TRY(sphneib.UpdateAllSphProperties(false));
  GDF_TRY(sphneib.BuildTree(true, 0, 1, 1, 0.0));
  GDF_TRY(sphneib.UpdateAllSphProperties(false));
  // a relative opening criterion needs forces to start from: the first pass uses the geometric
  // one, then the tree is built again (stocking those forces) and the forces recomputed without
  // a new h-rho pass (SphSimulation.cpp:380-470)
  const int mac = c->phys.gravity_mac;
  const bool twice = c->phys.self_gravity && mac != GDF_MAC_GEOMETRIC;
  if (twice) c->phys.gravity_mac = GDF_MAC_GEOMETRIC;
  const int frc = ComputeForces();
  c->phys.gravity_mac = mac;
  GDF_TRY(frc);
  if (twice) {
    GDF_TRY(sphneib.BuildTree(true, 0, 1, 1, 0.0));
    GDF_TRY(ComputeForces());
  }
  c->t = t0;
  c->tlast = t0;
  GDF_TRY(star_forces(c, false));      // the initial N-body forces (SphSimulation.cpp:490-502)
  if (c->nlevels > 1) {
    // the first ComputeBlockTimesteps resynchronises (n = nresync = 0); EndTimestep with the
    // particles' initial dt = 0 sets r0, v0, a0, u0, dudt0 and each particle's first step
    GDF_TRY(gdf::block_timesteps(c));
    GDF_TRY(gdf::block_end_step(c));
  } else {
    GDF_TRY(ComputeGlobalTimestep(&c->timestep));
    // EndTimestep with the particles' initial dt = 0 sets r0, v0, a0, u0, dudt0
    GDF_TRY(hydroint.EndTimestep(0, c->t, 0.0));
  }
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  c->nsteps = nsteps0;
  c->setup_done = true;
  c->rebuild_tree = true;          // the first main loop rebuilds (SphSimulation.cpp:339)
  return GDF_OK;
}

template <int ndim>
int GradhSphSimulationGpu<ndim>::MainLoop() {
  gdf_ctx* c = ctx;
  if (!c->setup_done) return gdf::fail(c, GDF_ERR_INVALID, "MainLoop before setup");
  if (c->nlevels > 1) return MainLoopBlock();
  GDF_TRY(gdf::zero_icount(c));
  record(c, 0);
  c->pass_ms[5] = 0.0;
  c->grav_launches = 0;
  c->pass_ms[6] = 0.0;
  c->hydro_launches = 0;
  c->nkev = 0;                 // kernel timings of earlier passes belong to no step
  c->nsteps++;
  const double dtstep = c->timestep;
  c->t = c->t + c->timestep;
  GDF_TRY(hydroint.AdvanceParticles((int) c->nsteps, c->t, c->timestep));
  GDF_TRY(hydroint.CheckBoundaries());
  record(c, 1);
  GDF_TRY(sphneib.BuildTree(c->rebuild_tree, (int) c->nsteps, c->cfg.ntreebuildstep,
                            c->cfg.ntreestockstep, c->timestep));
  c->rebuild_tree = false;
  record(c, 2);
  GDF_TRY(sphneib.UpdateAllSphProperties(false));
  record(c, 5);
  GDF_TRY(ComputeForces());
  GDF_TRY(star_forces(c, true));
  record(c, 4);
  double dtnew = 0.0;
  GDF_TRY(ComputeGlobalTimestep(&dtnew));
  GDF_TRY(hydroint.EndTimestep((int) c->nsteps, c->t, dtstep));
  record(c, 6);
  c->timestep = dtnew;
  GDF_HIP(c, hipMemcpyAsync(c->h_red + 16000, c->icount, sizeof(long long)*6,
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  const long long* ic = reinterpret_cast<const long long*>(c->h_red + 16000);
  for (int k = 0; k < 6; k++) c->icounts_last[k] = ic[k];
  c->finished_last = c->N;
  FinishStepStats();
  return GDF_OK;
}

template <int ndim>
void GradhSphSimulationGpu<ndim>::FinishStepStats() {
  gdf_ctx* c = ctx;
  if (c->timing) {
    gdf::ktime_collect(c);
    float ms = 0.f;
    // tree, density, hydro, gravity, integrate
    hipEventElapsedTime(&ms, c->ev[1], c->ev[2]); c->pass_ms[0] = ms;
    hipEventElapsedTime(&ms, c->ev[2], c->ev[5]); c->pass_ms[1] = ms;
    hipEventElapsedTime(&ms, c->ev[5], c->ev[3]); c->pass_ms[2] = ms;
    hipEventElapsedTime(&ms, c->ev[3], c->ev[4]); c->pass_ms[3] = ms;
    float a = 0.f, b = 0.f;
    hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
    hipEventElapsedTime(&b, c->ev[4], c->ev[6]);
    c->pass_ms[4] = a + b;
  }
}

template <int ndim>
int GradhSphSimulationGpu<ndim>::MainLoopBlock() {
  gdf_ctx* c = ctx;
  GDF_TRY(gdf::zero_icount(c));
  record(c, 0);
  c->pass_ms[5] = 0.0;
  c->grav_launches = 0;
  c->pass_ms[6] = 0.0;
  c->hydro_launches = 0;
  c->nkev = 0;                 // kernel timings of earlier passes belong to no step
  c->n_int += 1;
  c->nsteps++;
  c->t = c->t + c->timestep;
  GDF_TRY(gdf::block_advance(c));
  GDF_TRY(hydroint.CheckBoundaries());
  record(c, 1);
  GDF_TRY(sphneib.BuildTree(c->rebuild_tree, (int) c->nsteps, c->cfg.ntreebuildstep,
                            c->cfg.ntreestockstep, c->timestep));
  c->rebuild_tree = false;
  record(c, 2);
  // the tree build synchronised the stream: block_advance's count of active particles is in
  c->n_active = c->h_counters[12];
  int activecount = 0, rounds = 0;
  do {
    c->mark_passes = rounds == 0;
    GDF_TRY(sphneib.UpdateAllSphProperties(false));
    record(c, 5);
    GDF_TRY(ComputeForces());
    record(c, 4);
    c->mark_passes = true;
    // SphLeapfrogKDK::CheckTimesteps: the active flags are rewritten, set only for the
    // particles whose step a neighbour's level ends early
    GDF_TRY(gdf::block_check_timesteps(c, &activecount));
    c->n_active = activecount;
    if (++rounds > 64) return gdf::fail(c, GDF_ERR_TIMESTEP, "CheckTimesteps did not settle");
  } while (activecount > 0);
  GDF_TRY(gdf::block_timesteps(c));
  GDF_TRY(gdf::block_end_step(c));
  c->tlast = c->t;
  record(c, 6);
  GDF_HIP(c, hipMemcpyAsync(c->h_red + 16000, c->icount, sizeof(long long)*7,
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  const long long* ic = reinterpret_cast<const long long*>(c->h_red + 16000);
  for (int k = 0; k < 6; k++) c->icounts_last[k] = ic[k];
  c->finished_last = ic[6];
  FinishStepStats();
  return GDF_OK;
}

template <int ndim>
int NbodySimulationGpu<ndim>::ComputeForces() {
  gdf_ctx* c = ctx;
  GDF_TRY(gdf::integ_zero_forces(c));
  record(c, 3);
  if (c->cfg.nbody_direct) {
    GDF_TRY(gdf::nbody_direct(c, c->cfg.nbody_direct == 2));
    record(c, 4);
    return GDF_OK;
  }
  GDF_TRY(gdf::tree_build(c));
  record(c, 5);
  GDF_TRY(gdf::gravity_walk(c));
  GDF_TRY(gdf::integ_add_atree(c));
  record(c, 4);
  return GDF_OK;
}

template <int ndim>
int NbodySimulationGpu<ndim>::PostInitialConditionsSetup() {
  gdf_ctx* c = ctx;
  double t0;
  int64_t nsteps0;
  initial_state(c, &t0, &nsteps0);
  c->t = t0;
  c->tlast = t0;
  GDF_TRY(ComputeForces());
  GDF_TRY(gdf::integ_timestep(c, &c->timestep));
  GDF_TRY(nbodyint.EndTimestep(0, c->t, 0.0));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  c->nsteps = nsteps0;
  c->setup_done = true;
  return GDF_OK;
}

template <int ndim>
int NbodySimulationGpu<ndim>::MainLoop() {
  gdf_ctx* c = ctx;
  if (!c->setup_done) return gdf::fail(c, GDF_ERR_INVALID, "MainLoop before setup");
  GDF_TRY(gdf::zero_icount(c));
  record(c, 0);
  c->pass_ms[5] = 0.0;
  c->grav_launches = 0;
  c->pass_ms[6] = 0.0;
  c->hydro_launches = 0;
  c->nkev = 0;                 // kernel timings of earlier passes belong to no step
  c->nsteps++;
  const double dtstep = c->timestep;
  c->t = c->t + c->timestep;
  GDF_TRY(nbodyint.AdvanceParticles((int) c->nsteps, c->t, c->timestep));
  record(c, 1);
  GDF_TRY(ComputeForces());
  double dtnew = 0.0;
  GDF_TRY(gdf::integ_timestep(c, &dtnew));
  // NbodyLeapfrogKDK::CorrectionTerms and EndTimestep
  GDF_TRY(nbodyint.EndTimestep((int) c->nsteps, c->t, dtstep));
  record(c, 6);
  c->timestep = dtnew;
  GDF_HIP(c, hipMemcpyAsync(c->h_red + 16000, c->icount, sizeof(long long)*6,
                            hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  const long long* ic = reinterpret_cast<const long long*>(c->h_red + 16000);
  for (int k = 0; k < 6; k++) c->icounts_last[k] = ic[k];
  if (c->timing) {
    float a = 0.f, b = 0.f, f = 0.f, g = 0.f;
    for (int k = 0; k < 6; k++) c->pass_ms[k] = 0.0;
    hipEventElapsedTime(&a, c->ev[0], c->ev[1]);
    hipEventElapsedTime(&f, c->ev[3], c->ev[4]);
    hipEventElapsedTime(&b, c->ev[4], c->ev[6]);
    c->pass_ms[3] = f;
    c->pass_ms[4] = a + b;
    if (!c->cfg.nbody_direct) {
      hipEventElapsedTime(&g, c->ev[3], c->ev[5]);
      c->pass_ms[0] = g;
      c->pass_ms[3] = f - g;
    }
  }
  return GDF_OK;
}

template class NbodySimulationGpu<3>;
template class GradhSphTreeGpu<3>;
template class SphLeapfrogKDKGpu<3>;
template class GradhSphSimulationGpu<3>;

}  // namespace gandalf_gpu

extern "C" {

// the whole-step drivers integrate stars with one global timestep only (the per-method entry
// points serve a caller's own block-timestep loop)
static int stars_need_global_steps(gdf_ctx* c) {
  if (c->Nstar > 0 && c->nlevels > 1)
    return gdf::fail(c, GDF_ERR_INVALID, "stars beside the gas need one timestep level "
                                         "(Nlevels = 1) in gdf_setup_simulation / gdf_main_loop");
  return GDF_OK;
}

int gdf_setup_simulation(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  GDF_TRY(stars_need_global_steps(c));
  if (c->phys.nbody) {
    gandalf_gpu::NbodySimulationGpu<3> sim(c);
    return sim.PostInitialConditionsSetup();
  }
  gandalf_gpu::GradhSphSimulationGpu<3> sim(c);
  return sim.PostInitialConditionsSetup();
}

int gdf_main_loop(gdf_ctx* c, int nsteps) {
  if (!c || c->N <= 0 || nsteps < 0) return GDF_ERR_INVALID;
  GDF_TRY(stars_need_global_steps(c));
  if (c->phys.nbody) {
    gandalf_gpu::NbodySimulationGpu<3> sim(c);
    for (int s = 0; s < nsteps; s++) GDF_TRY(sim.MainLoop());
    return GDF_OK;
  }
  gandalf_gpu::GradhSphSimulationGpu<3> sim(c);
  for (int s = 0; s < nsteps; s++) GDF_TRY(sim.MainLoop());
  return GDF_OK;
}

int gdf_nbody_direct_forces(gdf_ctx* c, int smoothed) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  GDF_TRY(gdf::nbody_direct(c, smoothed ? 1 : 0));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

}  // extern "C"
