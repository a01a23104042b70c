// Reference-side binding of the device path: the two classes a GANDALF build adds to select it
// with "neib_search = gpu" and "sph_integration = lfkdk_gpu" (GradhSphSimulation::
// ProcessSphParameters, src/GradhSph/GradhSphSimulation.cpp:171-186, 232-246).
//
//   GradhSphTreeGpu<ndim, ParticleType>   : SphNeighbourSearch<ndim>
//                                           (src/Headers/SphNeighbourSearch.h:77-94,
//                                            src/Headers/NeighbourSearch.h:66-137)
//   SphLeapfrogKDKGpu<ndim, ParticleType> : SphIntegration<ndim>
//                                           (src/Headers/Integration.h:48-95)
//
// Both forward to the C ABI of libgandalf_gpu (include/gandalf_gpu.h) on one shared device
// context.  The device owns the particle state; the reference's particle array (AoS, in the
// reference's order) is the host mirror: positions, velocities, masses and internal energies go
// up before the tree build, and every pass brings back the fields the reference's own code
// reads next (SphSimulation::MainLoop, src/Hydrodynamics/SphSimulation.cpp:574-880).  Non-zero
// statuses become ExceptionHandler errors with the library's message.  Periodic images are taken
// on the device inside every pair test, so the ghost searches are empty and the ghost tree is
// never built; the MPI-only virtuals are not declared (the build is serial).
//
// This header compiles against the reference's own headers only (no reference source is
// needed here); tests/test_reference_binding.py instantiates both classes with them.
#ifndef GANDALF_GPU_REFERENCE_BINDING_H
#define GANDALF_GPU_REFERENCE_BINDING_H

#include <string>
#include <vector>

#include "Exception.h"
#include "Integration.h"
#include "Nbody.h"
#include "SphNeighbourSearch.h"
#include "gandalf_gpu.h"

namespace gandalf_gpu_binding {

inline void check(gdf_ctx* ctx, int rc) {
  if (rc < 0) ExceptionHandler::getIstance().raise(std::string(gdf_last_error(ctx)));
}

// The device context and the particle traffic between it and the reference's particle array.
template <int ndim, template <int> class ParticleType>
class DeviceState {
 public:
  explicit DeviceState(const gdf_config& cfg) : ctx(nullptr), N(0), Nstar(0) {
    const int rc = gdf_create(&cfg, &ctx);
    if (rc != GDF_OK) ExceptionHandler::getIstance().raise("gdf_create failed");
  }
  ~DeviceState() { gdf_destroy(ctx); }

  // r, v, m, h, u of the reference's particles to the device (all of them at the first call or
  // when the particle count changed, else the fields its host code may have changed)
  void Upload(Hydrodynamics<ndim>* hydro) {
    const int n = hydro->Nhydro;
    std::vector<double> r(3*n), v(3*n), m(n), h(n), u(n);
    for (int i = 0; i < n; i++) {
      const ParticleType<ndim>& p = Part(hydro, i);
      for (int k = 0; k < 3; k++) {
        r[3*i + k] = k < ndim ? p.r[k] : 0.0;
        v[3*i + k] = k < ndim ? p.v[k] : 0.0;
      }
      m[i] = p.m; h[i] = p.h; u[i] = p.u;
    }
    if (n != N) {
      check(ctx, gdf_set_particles(ctx, n, &r[0], &v[0], &m[0], &h[0], &u[0]));
      N = n;
      return;
    }
    check(ctx, gdf_set_field(ctx, "r", &r[0]));
    check(ctx, gdf_set_field(ctx, "v", &v[0]));
    check(ctx, gdf_set_field(ctx, "h", &h[0]));
    check(ctx, gdf_set_field(ctx, "u", &u[0]));
  }

  // the stars' r, v, m, h (Nbody<ndim>::stardata) to the device before a pass that reads them:
  // the h-rho pass (their zeta terms), the force passes (their pull on the gas) and the star-gas
  // sums
  void UploadStars(Nbody<ndim>* nbody) {
    const int n = nbody ? nbody->Nstar : 0;
    if (n == 0 && Nstar == 0) return;
    std::vector<double> r(3*n + 1, 0.0), v(3*n + 1, 0.0), m(n + 1), h(n + 1);
    for (int i = 0; i < n; i++) {
      const NbodyParticle<ndim>& st = nbody->stardata[i];
      for (int k = 0; k < ndim; k++) { r[3*i + k] = st.r[k]; v[3*i + k] = st.v[k]; }
      m[i] = st.m; h[i] = st.h;
    }
    check(ctx, gdf_set_stars(ctx, n, &r[0], &v[0], &m[0], &h[0]));
    Nstar = n;
  }

  // a device field into the reference's particles: width 1 or 3 (vectors by component)
  template <typename Set>
  void Download(Hydrodynamics<ndim>* hydro, const char* name, int width, Set set) {
    const int n = hydro->Nhydro;
    buf.resize((size_t) width*n);
    check(ctx, gdf_get_field(ctx, name, &buf[0]));
    for (int i = 0; i < n; i++) set(Part(hydro, i), &buf[(size_t) width*i]);
  }

  static ParticleType<ndim>& Part(Hydrodynamics<ndim>* hydro, int i) {
    return static_cast<ParticleType<ndim>&>(hydro->GetParticlePointer(i));
  }

  gdf_ctx* ctx;
  int N, Nstar;
  std::vector<double> buf;
};

template <int ndim, template <int> class ParticleType>
class GradhSphTreeGpu : public SphNeighbourSearch<ndim> {
 public:
  typedef DeviceState<ndim, ParticleType> State;

  GradhSphTreeGpu(State* state, MAC_Type mac) : dev(state), mac_type(geometric), timing(0),
                                                 neibcheck(false) {
    SetOpeningCriterion(mac);
  }
  virtual ~GradhSphTreeGpu() {}

  // ---- NeighbourSearch<ndim> ----
  // HydroTree::BuildTree (src/Tree/HydroTree.cpp:310-372) with its schedule: rebuild, or
  // refit of the kept tree on the other steps
  virtual void BuildTree(const bool rebuild_tree, const int n, const int ntreebuildstep,
                         const int ntreestockstep, const FLOAT, Hydrodynamics<ndim>* hydro) {
    dev->Upload(hydro);
    check(dev->ctx, gdf_build_tree_step(dev->ctx, rebuild_tree ? 1 : 0, n, ntreebuildstep,
                                        ntreestockstep));
  }
  virtual void BuildGhostTree(const bool, const int, const int, const int, const FLOAT,
                              Hydrodynamics<ndim>*) {}
  // HydroTree::GetGatherNeighbourList (HydroTree.cpp:451-471): -1 when Nneibmax is too small
  virtual int GetGatherNeighbourList(FLOAT* rp, FLOAT rsearch, Particle<ndim>*, int,
                                     int Nneibmax, int* neiblist) {
    double r[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < ndim; k++) r[k] = rp[k];
    const int n = gdf_get_gather_neighbour_list(dev->ctx, r, rsearch, Nneibmax, neiblist);
    if (n < -1) check(dev->ctx, n);
    return n;
  }
  // HydroTree::SearchBoundaryGhostParticles (HydroTree.cpp:495-543): no ghost copies
  virtual void SearchBoundaryGhostParticles(FLOAT, const DomainBox<ndim>&,
                                            Hydrodynamics<ndim>* hydro) {
    hydro->Nghost = 0;
    hydro->NPeriodicGhost = 0;
  }
  // HydroTree::UpdateActiveParticleCounters: the reference's active flags to the device
  virtual void UpdateActiveParticleCounters(Hydrodynamics<ndim>* hydro) {
    const int n = hydro->Nhydro;
    std::vector<double> act(n);
    for (int i = 0; i < n; i++) act[i] = State::Part(hydro, i).flags.check(active) ? 1.0 : 0.0;
    check(dev->ctx, gdf_set_field(dev->ctx, "active", &act[0]));
  }
  // HydroTree::UpdateAllStarGasForces (HydroTree.cpp:552-654): the gas's pull on every active
  // star, added to its a and gpot, through the brute-force lists (every gas particle direct)
  virtual void UpdateAllStarGasForces(Hydrodynamics<ndim>*, Nbody<ndim>* nbody,
                                      DomainBox<ndim>& simbox, Ewald<ndim>*) {
    if (simbox.PeriodicGravity)
      ExceptionHandler::getIstance().raise("neib_search = gpu: periodic gravity is not built");
    dev->UploadStars(nbody);
    check(dev->ctx, gdf_update_all_star_gas_forces(dev->ctx));
    const int n = nbody->Nstar;
    std::vector<double> a(3*n + 1), gpot(n + 1);
    check(dev->ctx, gdf_get_star_field(dev->ctx, "a", &a[0]));
    check(dev->ctx, gdf_get_star_field(dev->ctx, "gpot", &gpot[0]));
    for (int i = 0; i < n; i++) {
      NbodyParticle<ndim>& st = nbody->stardata[i];
      if (!st.flags.check(active)) continue;
      for (int k = 0; k < ndim; k++) st.a[k] += a[3*i + k];
      st.gpot += gpot[i];
    }
  }
  virtual double GetMaximumSmoothingLength() const {
    double hmax = 0.0;
    check(dev->ctx, gdf_maximum_smoothing_length(dev->ctx, &hmax));
    return hmax;
  }
  virtual TreeBase<ndim>* GetTree() const { return 0; }
  virtual TreeBase<ndim>* GetGhostTree() const { return 0; }
  virtual void SetTimingObject(CodeTiming* t) { timing = t; }
  virtual void ToggleNeighbourCheck(bool do_check) { neibcheck = do_check; }
  // used by the meshless finite-volume schemes only (MeshlessFVSimulation.cpp:745-748)
  virtual void UpdateTimestepsLimitsFromDistantParticles(Hydrodynamics<ndim>*, const bool) {
    ExceptionHandler::getIstance().raise("neib_search = gpu: grad-h SPH only");
  }
  // HydroTree::Get/SetOpeningCriterion (src/Tree/HydroTree.cpp:196-205); the reference's
  // MAC_Type (Tree.h:52-56) numbers the criteria geometric, eigenmac, gadget2
  virtual MAC_Type GetOpeningCriterion() const { return mac_type; }
  virtual void SetOpeningCriterion(MAC_Type mac) {
    const int dm = mac == eigenmac ? GDF_MAC_EIGENMAC
                 : mac == gadget2 ? GDF_MAC_GADGET2 : GDF_MAC_GEOMETRIC;
    check(dev->ctx, gdf_set_opening_criterion(dev->ctx, dm));
    mac_type = mac;
  }

  // ---- SphNeighbourSearch<ndim> ----
  // GradhSphTree::UpdateAllSphProperties (src/GradhSph/GradhSphTree.cpp:83-271)
  virtual void UpdateAllSphProperties(Sph<ndim>* sph, Nbody<ndim>* nbody) {
    dev->UploadStars(nbody);
    check(dev->ctx, gdf_update_all_sph_properties(dev->ctx, 0));
    const FLOAT krsqd = sph->kernp->kernrangesqd;
    dev->Download(sph, "h", 1, [krsqd](ParticleType<ndim>& p, const double* x) {
      p.h = x[0]; p.hrangesqd = krsqd*p.h*p.h; });
    dev->Download(sph, "rho", 1, [](ParticleType<ndim>& p, const double* x) { p.rho = x[0]; });
    dev->Download(sph, "invomega", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.invomega = x[0]; });
    dev->Download(sph, "zeta", 1, [](ParticleType<ndim>& p, const double* x) { p.zeta = x[0]; });
    dev->Download(sph, "hfactor", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.hfactor = x[0]; });
    dev->Download(sph, "sound", 1, [](ParticleType<ndim>& p, const double* x) { p.sound = x[0]; });
    dev->Download(sph, "pressure", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.pressure = x[0]; });
    dev->Download(sph, "u", 1, [](ParticleType<ndim>& p, const double* x) { p.u = x[0]; });
  }
  // GradhSphTree::UpdateAllSphHydroForces (GradhSphTree.cpp:280-435)
  virtual void UpdateAllSphHydroForces(Sph<ndim>* sph, Nbody<ndim>* nbody, DomainBox<ndim>&) {
    dev->UploadStars(nbody);
    check(dev->ctx, gdf_update_all_sph_hydro_forces(dev->ctx));
    DownloadForces(sph, nbody && nbody->Nstar > 0);
  }
  // GradhSphTree::UpdateAllSphForces (GradhSphTree.cpp:444-657); periodic gravity with the
  // device's own Ewald table when all three axes are periodic (the device context must be
  // created with periodic = 1 and self_gravity = 1)
  virtual void UpdateAllSphForces(Sph<ndim>* sph, Nbody<ndim>* nbody, DomainBox<ndim>& simbox,
                                  Ewald<ndim>*) {
    if (simbox.PeriodicGravity)
      for (int k = 0; k < 3; k++)
        if (simbox.boundary_lhs[k] != periodicBoundary ||
            simbox.boundary_rhs[k] != periodicBoundary)
          ExceptionHandler::getIstance().raise("neib_search = gpu: periodic gravity needs all "
                                               "three axes periodic");
    dev->UploadStars(nbody);
    check(dev->ctx, gdf_update_all_sph_forces(dev->ctx));
    DownloadForces(sph, true);
  }

 private:
  void DownloadForces(Sph<ndim>* sph, bool grav) {
    dev->Download(sph, "a", 3, [](ParticleType<ndim>& p, const double* x) {
      for (int k = 0; k < ndim; k++) p.a[k] = x[k]; });
    if (grav) {
      dev->Download(sph, "atree", 3, [](ParticleType<ndim>& p, const double* x) {
        for (int k = 0; k < ndim; k++) p.atree[k] = x[k]; });
      dev->Download(sph, "gpot", 1, [](ParticleType<ndim>& p, const double* x) { p.gpot = x[0]; });
    }
    dev->Download(sph, "dudt", 1, [](ParticleType<ndim>& p, const double* x) { p.dudt = x[0]; });
    dev->Download(sph, "div_v", 1, [](ParticleType<ndim>& p, const double* x) { p.div_v = x[0]; });
    dev->Download(sph, "levelneib", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.levelneib = (int) x[0]; });
  }

  State* dev;
  MAC_Type mac_type;
  CodeTiming* timing;
  bool neibcheck;
};

template <int ndim, template <int> class ParticleType>
class SphLeapfrogKDKGpu : public SphIntegration<ndim> {
 public:
  typedef DeviceState<ndim, ParticleType> State;

  SphLeapfrogKDKGpu(State* state, DOUBLE accel_mult, DOUBLE courant_mult, DOUBLE energy_mult,
                    bool energy_integration, tdaviscenum tdavisc)
    : SphIntegration<ndim>(accel_mult, courant_mult, energy_mult, energy_integration, tdavisc),
      dev(state) {}
  virtual ~SphLeapfrogKDKGpu() {}

  // SphLeapfrogKDK::AdvanceParticles (SphLeapfrogKDK.cpp:76-127) on the device
  void AdvanceParticles(const int n, const FLOAT t, const FLOAT timestep,
                        Hydrodynamics<ndim>* hydro) {
    dev->Upload(hydro);
    check(dev->ctx, gdf_advance_particles_block(dev->ctx, n, t, timestep));
    DownloadState(hydro);
    DownloadActive(hydro);
  }
  // SphLeapfrogKDK::CorrectionTerms (SphLeapfrogKDK.cpp:138-183)
  void CorrectionTerms(const int n, const FLOAT t, const FLOAT timestep,
                       Hydrodynamics<ndim>* hydro) {
    check(dev->ctx, gdf_correction_terms(dev->ctx, n, t, timestep));
    DownloadState(hydro);
  }
  // SphLeapfrogKDK::SetActiveParticles (SphLeapfrogKDK.cpp:192-211)
  void SetActiveParticles(const int n, Hydrodynamics<ndim>* hydro) {
    check(dev->ctx, gdf_set_active_particles(dev->ctx, n));
    DownloadActive(hydro);
  }
  // SphLeapfrogKDK::EndTimestep (SphLeapfrogKDK.cpp:219-272)
  void EndTimestep(const int n, const FLOAT t, const FLOAT timestep,
                   Hydrodynamics<ndim>* hydro) {
    check(dev->ctx, gdf_end_timestep_block(dev->ctx, n, t, timestep));
    DownloadState(hydro);
    dev->Download(hydro, "nlast", 1, [](ParticleType<ndim>& p, const double* x) {
      p.nlast = (int) x[0]; p.flags.unset(active); p.flags.unset(end_timestep); });
    dev->Download(hydro, "tlast", 1, [](ParticleType<ndim>& p, const double* x) { p.tlast = x[0]; });
  }
  // SphLeapfrogKDK::CheckTimesteps (SphLeapfrogKDK.cpp:284-329)
  int CheckTimesteps(const int level_diff_max, const int level_step, const int n,
                     const FLOAT timestep, Hydrodynamics<ndim>* hydro) {
    int activecount = 0;
    check(dev->ctx, gdf_check_timesteps(dev->ctx, level_diff_max, level_step, n, timestep,
                                        &activecount));
    dev->Download(hydro, "level", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.level = (int) x[0]; });
    dev->Download(hydro, "nstep", 1,
                  [](ParticleType<ndim>& p, const double* x) { p.nstep = (int) x[0]; });
    DownloadActive(hydro);
    return activecount;
  }

 private:
  void DownloadState(Hydrodynamics<ndim>* hydro) {
    dev->Download(hydro, "r", 3, [](ParticleType<ndim>& p, const double* x) {
      for (int k = 0; k < ndim; k++) p.r[k] = x[k]; });
    dev->Download(hydro, "v", 3, [](ParticleType<ndim>& p, const double* x) {
      for (int k = 0; k < ndim; k++) p.v[k] = x[k]; });
    dev->Download(hydro, "u", 1, [](ParticleType<ndim>& p, const double* x) { p.u = x[0]; });
  }
  void DownloadActive(Hydrodynamics<ndim>* hydro) {
    dev->Download(hydro, "active", 1, [](ParticleType<ndim>& p, const double* x) {
      if (x[0] != 0.0) p.flags.set(active); else p.flags.unset(active); });
  }

  State* dev;
};

}  // namespace gandalf_gpu_binding

#endif
