// Host-side mirror of the reference classes that drive the particle hot path, backed by the
// device passes of libgandalf_gpu.  Method names and call order follow the reference so a
// MainLoop-shaped caller reads the same:
//
//   GradhSphTreeGpu      ~ SphNeighbourSearch<ndim> / GradhSphTree (src/Headers/SphNeighbourSearch.h:77-94)
//   SphLeapfrogKDKGpu    ~ TimeIntegration<ndim> / SphLeapfrogKDK (src/Headers/Integration.h:48-66)
//   GradhSphSimulationGpu~ SphSimulation<ndim> (src/Hydrodynamics/SphSimulation.cpp:204-880)
//   NbodySimulationGpu   ~ NbodySimulation<ndim> (src/Nbody/NbodySimulation.cpp:183-375)
//
// Each call returns a GDF status code instead of raising; the C ABI in gandalf_gpu.h exposes
// the simulation driver as gdf_setup_simulation / gdf_main_loop.
#pragma once

#include "../gdf_internal.h"

namespace gandalf_gpu {

template <int ndim>
class GradhSphTreeGpu {
 public:
  explicit GradhSphTreeGpu(gdf_ctx* c) : ctx(c) {}
  // HydroTree::BuildTree (HydroTree.cpp:310-372): gdf_build_tree_step's schedule
  int BuildTree(bool rebuild_tree, int n, int ntreebuildstep, int ntreestockstep,
                double timestep);
  // GradhSphTree::UpdateAllSphProperties (GradhSphTree.cpp:83-271)
  int UpdateAllSphProperties(bool nopbc);
  // GradhSphTree::UpdateAllSphHydroForces (GradhSphTree.cpp:280-435)
  int UpdateAllSphHydroForces();
  // GradhSphTree::UpdateAllSphForces (GradhSphTree.cpp:444-657)
  int UpdateAllSphForces();
  // HydroTree::GetMaximumSmoothingLength
  int GetMaximumSmoothingLength(double* hmax);

 private:
  gdf_ctx* ctx;
};

template <int ndim>
class SphLeapfrogKDKGpu {
 public:
  explicit SphLeapfrogKDKGpu(gdf_ctx* c) : ctx(c) {}
  // SphLeapfrogKDK::AdvanceParticles (SphLeapfrogKDK.cpp:76-127): dt = t - tlast
  int AdvanceParticles(int n, double t, double timestep);
  // TimeIntegration::CheckBoundaries (Integration.cpp:11-80)
  int CheckBoundaries();
  // SphLeapfrogKDK::EndTimestep (SphLeapfrogKDK.cpp:219-272)
  int EndTimestep(int n, double t, double timestep);

 private:
  gdf_ctx* ctx;
};

template <int ndim>
class GradhSphSimulationGpu {
 public:
  explicit GradhSphSimulationGpu(gdf_ctx* c) : ctx(c), sphneib(c), hydroint(c) {}
  // SphSimulation::PostInitialConditionsSetup (SphSimulation.cpp:204-565), file ICs
  int PostInitialConditionsSetup();
  // SphSimulation::MainLoop (SphSimulation.cpp:574-880): global timesteps for Nlevels = 1,
  // block timesteps otherwise
  int MainLoop();
  // Simulation::ComputeGlobalTimestep (Simulation.cpp:1669-1754)
  int ComputeGlobalTimestep(double* dt);

 private:
  int ComputeForces();
  // MainLoop with block timesteps: per-particle drift, forces of the active particles, the
  // CheckTimesteps loop, Simulation::ComputeBlockTimesteps (Simulation.cpp:1764-2197) and
  // EndTimestep of the particles whose step ended
  int MainLoopBlock();
  void FinishStepStats();
  gdf_ctx* ctx;
  GradhSphTreeGpu<ndim> sphneib;
  SphLeapfrogKDKGpu<ndim> hydroint;
};

// Pure N-body run: stars only.  Forces by the reference's direct sums (nbody_direct 1/2) or by
// the tree walk with Plummer softening; NbodyLeapfrogKDK's AdvanceParticles / CorrectionTerms /
// EndTimestep (src/Nbody/NbodyLeapfrogKDK.cpp:250-370) map onto the same device KDK kernels as
// the gas (no internal energy).
template <int ndim>
class NbodySimulationGpu {
 public:
  explicit NbodySimulationGpu(gdf_ctx* c) : ctx(c), nbodyint(c) {}
  // NbodySimulation::PostInitialConditionsSetup (NbodySimulation.cpp:183-250)
  int PostInitialConditionsSetup();
  // NbodySimulation::MainLoop (NbodySimulation.cpp:258-375) with Nlevels = 1
  int MainLoop();
 private:
  int ComputeForces();
  gdf_ctx* ctx;
  SphLeapfrogKDKGpu<ndim> nbodyint;
};

}  // namespace gandalf_gpu
