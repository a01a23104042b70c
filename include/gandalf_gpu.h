/* gandalf_gpu.h -- C ABI of the MI355X particle hot path (grad-h SPH + tree gravity).
 *
 * Plain pointers and sizes only.  Host arrays are in the caller's particle order; positions,
 * velocities and accelerations are (N,3) row-major doubles.  Every entry point returns
 * GDF_OK (0) or a negative status; gdf_last_error() gives the message.  A context owns one
 * HIP stream and its device memory and is not re-entrant; calls are ordered on the stream and
 * host-visible results are synchronised before return.
 *
 * Each entry point names the reference interface it replaces (gandalfcode/gandalf src/).
 */
#ifndef GANDALF_GPU_H
#define GANDALF_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  GDF_OK = 0,
  GDF_ERR_INVALID = -1,      /* invalid argument or call order */
  GDF_ERR_NOMEM = -2,        /* device allocation failed */
  GDF_ERR_HIP = -3,          /* HIP runtime error */
  GDF_ERR_NOCONV = -4,       /* h-rho iteration did not converge (GradhSph.cpp:245-251) */
  GDF_ERR_CAPACITY = -5,     /* neighbour/walk capacity exhausted after internal retries */
  GDF_ERR_TIMESTEP = -6      /* invalid timestep (SphIntegration.cpp:122-129) */
};

enum { GDF_KERNEL_M4 = 0, GDF_KERNEL_QUINTIC = 1 };
enum { GDF_EOS_ENERGY_EQN = 0, GDF_EOS_ISOTHERMAL = 1, GDF_EOS_BAROTROPIC = 2 };
enum { GDF_PREC_FP64 = 0, GDF_PREC_MIXED = 1 };

/* Run-time parameters, named after the reference's parameter keys
 * (src/Common/Parameters.cpp:160-575). */
typedef struct gdf_config {
  int kernel;                 /* "kernel": m4 | quintic */
  int tabulated_kernel;       /* "tabulated_kernel" (1000-entry tables) */
  double h_fac, h_converge;   /* "h_fac", "h_converge" */
  int gas_eos;                /* "gas_eos": energy_eqn | isothermal | barotropic */
  double gamma_eos, temp0, mu_bar, rho_bary;
  double alpha_visc, beta_visc;   /* "avisc" = mon97 */
  int hydro_forces, self_gravity;
  double thetamaxsqd;         /* geometric MAC opening angle squared */
  int Nleafmax;               /* leaf bucket size of the walks */
  int periodic;               /* all three axes periodic ("boundary_lhs/rhs" = periodic) */
  double boxmin[3], boxmax[3];
  double courant_mult, accel_mult, energy_mult;
  int precision;              /* GDF_PREC_FP64 or GDF_PREC_MIXED (FP32 pair arithmetic) */
  int group_max;              /* particles per walk group (<= 64) */
  /* pure N-body runs (NbodySimulation, src/Nbody/NbodySimulation.cpp): the particles are
   * stars, there is no gas pass; forces by the tree walk with Plummer softening nbody_eps, or
   * by the reference's direct sums (nbody_direct 1 = unsoftened, 2 = kernel-smoothed over h) */
  int nbody;
  int nbody_direct;
  double nbody_eps;
  double nbody_mult;          /* "nbody_mult" timestep multiplier */
  double fixed_dt;            /* > 0: every step uses this timestep */
  /* block timesteps (Simulation.cpp:1764-2197): "Nlevels" (1 = global timestep) and
   * "level_diff_max" (max. level difference to neighbours) */
  int Nlevels;
  int level_diff_max;
  /* failure-path controls, 0 = the reference's behaviour: the fixed-point iterations of the
   * h-rho iteration before it turns to bisection (GradhSph.cpp:156, iteration_max = 30; it
   * fails after five times as many), and the candidate batches of 64 reserved per walk group
   * before the first overflow regrows them (16) */
  int hrho_iteration_max;
  int cand_batches_init;
  /* restart (the reference's "gandalf -r", SimulationBase::restart): a snapshot read keeps its
   * snapshot counters (Noutsnap, Noutlitesnap, tsnaplast, tlitesnaplast) and the setup keeps
   * the file's particle ids; otherwise both are reset as for "ic = file"
   * (SimulationIO.hpp:1385-1394, SphSimulation.cpp:215-217) */
  int restart;
  /* stars beside the gas (gdf_set_stars): "nbody_softening" (1 = star-star sums and the stars'
   * zeta terms over the mean smoothing length of the pair, 0 = star-star point masses and the
   * gas particle's own h) and "conservative_sph_star_gravity" (the Hubber et al. 2013 zeta
   * terms of the stars in the gas h-rho pass, GradhSph.cpp:282-307) */
  int nbody_softening;
  int conservative_sph_star_gravity;
  /* Ewald periodic gravity, on whenever the box is periodic and self_gravity is set
   * (Simulation.cpp:1183-1198): "nEwaldGrid" table nodes per axis, "gr_bhewaldseriesn"
   * image / Fourier series radius, "ewald_mult" splitting factor, "EFratio" table extent in
   * half boxes (1 <= EFratio <= 1.5); Parameters.cpp:387-396 */
  int n_ewald_grid;
  int gr_bhewaldseriesn;
  double ewald_mult;
  double efratio;
  /* "ntreebuildstep", "ntreestockstep" (Parameters.cpp:311-312): the main loop rebuilds the tree
   * every ntreebuildstep steps and refits the kept tree to the particles' new positions on the
   * others (HydroTree::BuildTree, HydroTree.cpp:310-372) */
  int ntreebuildstep;
  int ntreestockstep;
  /* 1: the tree build's launches (bounds, keys, sort, permutation, topology, moments, group
   * selection) run as a HIP graph captured once per buffer arrangement and replayed; 0 (the
   * default): eager launches.  The results are the same bit for bit.  For a context driven by
   * one host thread with no other thread using the device meanwhile; serving pools refuse it. */
  int graph_capture;
  /* "multipole" (Parameters.cpp:309): the gravity walk's accepted-cell terms.
   * GDF_MULTIPOLE_QUADRUPOLE (the default) and GDF_MULTIPOLE_MONOPOLE evaluate every cell at
   * every target (ComputeCellQuadrupoleForces / ComputeCellMonopoleForces); the fast options sum
   * the cells once per walk group at the group's centre of mass and apply a first-order Taylor
   * series to each target (ComputeFastMonopoleForces / ComputeFastQuadrupoleForces,
   * GradhSphTree.cpp:587-593).  A cell is then accepted only beyond the group's extent about
   * that point divided by theta.  Gas walks only: the N-body tree walk keeps quadrupoles. */
  int multipole;
  /* "gravity_mac", "macerror" (Parameters.cpp:308, 314): the cell opening criterion of the
   * gravity walk (Tree::open_cell_for_gravity, Tree.h:413-432).  GDF_MAC_GEOMETRIC (the default)
   * accepts a cell beyond max(rmax, hmax)/theta; GDF_MAC_GADGET2 also opens a cell with
   * m rmax^2 / d^4 above macerror times the smallest |atree| of the walk group's particles,
   * GDF_MAC_EIGENMAC one nearer than (0.5 lambda / macerror)^(2/3) (lambda bounding its
   * quadrupole's largest eigenvalue) times the largest gpot^(-2/3) of the group's particles;
   * both read atree and gpot as they are at the tree build (KDTree.cpp:900-903, 1058-1079).
   * d is the distance from the walk group's box to the cell. */
  int gravity_mac;
  double macerror;
} gdf_config;

enum { GDF_MULTIPOLE_QUADRUPOLE = 0, GDF_MULTIPOLE_MONOPOLE = 1, GDF_MULTIPOLE_FAST_MONOPOLE = 2,
       GDF_MULTIPOLE_FAST_QUADRUPOLE = 3 };
enum { GDF_MAC_GEOMETRIC = 0, GDF_MAC_GADGET2 = 1, GDF_MAC_EIGENMAC = 2 };

typedef struct gdf_ctx gdf_ctx;

/* reference defaults (Parameters.cpp:160-575) for the fields above */
int gdf_config_default(gdf_config* cfg);

/* context lifetime; replaces GradhSphSimulation::ProcessSphParameters object creation
 * (src/GradhSph/GradhSphSimulation.cpp:54-260) for the hot path objects */
int gdf_create(const gdf_config* cfg, gdf_ctx** out);
void gdf_destroy(gdf_ctx* ctx);
const char* gdf_last_error(const gdf_ctx* ctx);

/* particle import; replaces Simulation::ImportArray (src/Common/Simulation.cpp:1442-1620).
 * h may be NULL (no initial smoothing lengths).  Resets all derived state. */
int gdf_set_particles(gdf_ctx* ctx, int N, const double* r, const double* v, const double* m,
                      const double* h, const double* u);
/* read a per-particle field back in import order; replaces the host particle array reads of
 * SphSnapshot::CopyDataFromSimulation (src/Common/SphSnapshot.cpp:340-372).  Names:
 * r v a atree r0 v0 a0 m h rho u u0 dudt dudt0 div_v gpot invomega zeta sound pressure hfactor
 * dt (timestep criterion) adot; block-timestep state: tlast dtstep (the particle's step
 * length, Particle.h dt) level levelneib nstep nlast active (integers returned as doubles) */
int gdf_get_field(gdf_ctx* ctx, const char* name, double* out);
/* write a per-particle field in import order (the names above; integers passed as doubles) */
int gdf_set_field(gdf_ctx* ctx, const char* name, const double* in);
int gdf_num_particles(const gdf_ctx* ctx);

/* ---- hot path: SphNeighbourSearch<3> (src/Headers/SphNeighbourSearch.h:77-94) ---- */
/* HydroTree::BuildTree (src/Tree/HydroTree.cpp:310-372): Morton sort, LBVH, moments */
int gdf_build_tree(gdf_ctx* ctx);
/* HydroTree::BuildTree(rebuild_tree, n, ntreebuildstep, ntreestockstep) with its schedule: a
 * full rebuild when rebuild_tree is set, n % ntreebuildstep == 0 or the particles changed since
 * the last build; otherwise the kept tree is refitted to the current positions, masses and
 * smoothing lengths (where the reference stocks the tree or extrapolates its cells, Tree.cpp:
 * 172-198: a refit keeps every node's box exact, which exact neighbour sets need).  Returns 1
 * after a rebuild, 0 after a refit, or a negative status. */
int gdf_build_tree_step(gdf_ctx* ctx, int rebuild_tree, int n, int ntreebuildstep,
                        int ntreestockstep);
/* NeighbourSearch::Get/SetOpeningCriterion (src/Headers/NeighbourSearch.h:99-100,
 * HydroTree.cpp:196-205): the gravity walk's criterion, GDF_MAC_*.  The gadget2 and eigenmac
 * criteria read each walk group's atree / gpot as stocked by the last tree build or refit, so a
 * caller switching to one builds the tree afterwards (as the reference's setup does,
 * SphSimulation.cpp:380-470).  The setter returns GDF_OK or GDF_ERR_INVALID; the getter the
 * criterion, or a negative status. */
int gdf_set_opening_criterion(gdf_ctx* ctx, int gravity_mac);
int gdf_opening_criterion(const gdf_ctx* ctx);
/* state of the captured tree-build graphs (cfg.graph_capture; no counterpart in the
 * reference): replays so far, and 1 if they are in use, 0 if capture is off, or minus the
 * reason a capture was given up (the context then builds eagerly): 1 capture could not begin,
 * 2 a launch failed, 3 capture could not end, 4 instantiation failed, 100 + the type of a
 * node that is neither a kernel nor a copy */
int gdf_graph_state(const gdf_ctx* ctx, int64_t* replays, int* in_use);
/* GradhSphTree::UpdateAllSphProperties (src/GradhSph/GradhSphTree.cpp:83-271);
 * nopbc = 1 ignores periodic images (the reference's first setup pass runs before its ghost
 * particles exist, SphSimulation.cpp:273-281) */
int gdf_update_all_sph_properties(gdf_ctx* ctx, int nopbc);
/* GradhSphTree::UpdateAllSphHydroForces (src/GradhSph/GradhSphTree.cpp:280-435) */
int gdf_update_all_sph_hydro_forces(gdf_ctx* ctx);
/* GradhSphTree::UpdateAllSphForces (src/GradhSph/GradhSphTree.cpp:444-657) */
int gdf_update_all_sph_forces(gdf_ctx* ctx);
/* HydroTree::GetMaximumSmoothingLength (src/Tree/HydroTree.cpp) */
int gdf_maximum_smoothing_length(gdf_ctx* ctx, double* hmax);

/* ---- TimeIntegration<3> (src/Headers/Integration.h:48-66), SphLeapfrogKDK ---- */
/* SphLeapfrogKDK::AdvanceParticles (src/Hydrodynamics/SphLeapfrogKDK.cpp:76-127);
 * dt = t - tlast as the reference computes it */
int gdf_advance_particles(gdf_ctx* ctx, double dt);
/* TimeIntegration::CheckBoundaries, periodic wrap (src/Common/Integration.cpp:11-80) */
int gdf_check_boundaries(gdf_ctx* ctx);
/* SphLeapfrogKDK::EndTimestep (src/Hydrodynamics/SphLeapfrogKDK.cpp:219-272) */
int gdf_end_timestep(gdf_ctx* ctx, double dtstep);
/* SphIntegration::Timestep + Simulation::ComputeGlobalTimestep
 * (SphIntegration.cpp:81-134, src/Common/Simulation.cpp:1669-1754) */
int gdf_compute_global_timestep(gdf_ctx* ctx, double* dt_min);
/* Simulation::CalculateDiagnostics (src/Common/SimAnalysis.hpp:52-185):
 * out = {Etot, ketot, utot, gpetot, mom[3], angmom[3], mtot} */
int gdf_calculate_diagnostics(gdf_ctx* ctx, double* out);

/* The per-method form of TimeIntegration<3> and of Simulation::ComputeBlockTimesteps, for a
 * caller that runs SphSimulation::MainLoop's sequence itself, block timesteps included
 * (SphSimulation.cpp:574-880).  n, t and timestep are the caller's integer time, time and base
 * timestep, as in the reference's signatures; each call sets the context's block clock from
 * them.  With one timestep level every particle is active at every integer time. */
/* SphLeapfrogKDK::AdvanceParticles(n, t, timestep) (SphLeapfrogKDK.cpp:76-127): each particle
 * drifts from the start of its own step (t - tlast) and is active iff its step ends at n */
int gdf_advance_particles_block(gdf_ctx* ctx, int n, double t, double timestep);
/* SphLeapfrogKDK::SetActiveParticles (SphLeapfrogKDK.cpp:192-211): active iff the particle's
 * end-of-step flag is set or n - nlast == nstep */
int gdf_set_active_particles(gdf_ctx* ctx, int n);
/* SphLeapfrogKDK::CorrectionTerms (SphLeapfrogKDK.cpp:138-183): particles whose step ends at n
 * take v += dt (a - a0)/2 (and the internal-energy correction with its first-order fallback) */
int gdf_correction_terms(gdf_ctx* ctx, int n, double t, double timestep);
/* SphLeapfrogKDK::CheckTimesteps (SphLeapfrogKDK.cpp:284-329): particles whose neighbours'
 * level exceeds theirs by more than level_diff_max end their step early where it is
 * synchronised; *activecount = their number (0 with one level) */
int gdf_check_timesteps(gdf_ctx* ctx, int level_diff_max, int level_step, int n,
                        double timestep, int* activecount);
/* Simulation::ComputeBlockTimesteps (src/Common/Simulation.cpp:1764-2197): new levels,
 * level_max / level_step / nresync / dt_max and the base timestep (read them back with
 * gdf_block_state and gdf_time: the integer time may be rescaled) */
int gdf_compute_block_timesteps(gdf_ctx* ctx);
/* SphLeapfrogKDK::EndTimestep(n, t, timestep) (SphLeapfrogKDK.cpp:219-272): particles whose
 * step ended close it (kick, r0/v0/a0, nlast = n, tlast = t) */
int gdf_end_timestep_block(gdf_ctx* ctx, int n, double t, double timestep);
/* NeighbourSearch::GetGatherNeighbourList (src/Headers/NeighbourSearch.h:83) ->
 * HydroTree::GetGatherNeighbourList (src/Tree/HydroTree.cpp:451-471, Tree.cpp:208-280): the
 * import-order ids, ascending, of the particles with |r - rp|^2 < rsearch^2, periodic images
 * included (the reference finds those through its ghost tree, as copies of these particles).
 * Returns their number, or -1 when it exceeds nmax (the reference's overflow code). */
int gdf_get_gather_neighbour_list(gdf_ctx* ctx, const double* rp, double rsearch, int nmax,
                                  int* list);

/* ---- whole-step drivers: SphSimulation<3> (src/Hydrodynamics/SphSimulation.cpp) ---- */
/* PostInitialConditionsSetup (:204-565) for initial conditions without smoothing lengths */
int gdf_setup_simulation(gdf_ctx* ctx);
/* MainLoop (:574-880), nsteps steps; global timesteps for Nlevels = 1, else block timesteps
 * (Simulation::ComputeBlockTimesteps, src/Common/Simulation.cpp:1764-2197) */
int gdf_main_loop(gdf_ctx* ctx, int nsteps);
/* block-timestep state (Simulation.h:187-215): iout = {n, level_max, level_step, nresync,
 * particles that ended a step in the last MainLoop}; dout = {dt_max} */
int gdf_block_state(const gdf_ctx* ctx, int64_t* iout5, double* dout1);
/* snapshots of the device state (Simulation::WriteSnapshotFile / ReadSnapshotFile,
 * src/Common/SimulationIO.hpp:60-150): format "column" (WriteColumnSnapshotFile /
 * ReadColumnSnapshotFile, :444-548, :193-272), "su" (WriteSerenUnformSnapshotFile /
 * ReadSerenUnformSnapshotFile, :2009-2249, :1294-1540) or "sf" (WriteSerenFormSnapshotFile /
 * ReadSerenFormSnapshotFile, :993-1222, :601-925); writing also "slite" / "seren_lite"
 * (WriteSerenLiteSnapshotFile, :2264-2497; the reference has no reader for it).  Reading sets
 * the particles (r, v, m, h, u; rho as a field) and the time, and "su" the step count; an
 * N-body context reads and writes the star block. */
int gdf_write_snapshot(gdf_ctx* ctx, const char* path, const char* format);
int gdf_read_snapshot(gdf_ctx* ctx, const char* path, const char* format);
/* the snapshot counters the Seren writers put in their headers and a restart read restores
 * (Simulation::Noutsnap, Noutlitesnap, tsnaplast, tlitesnaplast; SimulationIO.hpp:1385-1394,
 * 2145-2160): get2 = {Noutsnap, Noutlitesnap}, get2d = {tsnaplast, tlitesnaplast} receive the
 * context's values, then set2 / set2d (if not NULL) replace them.  The caller's output loop
 * (Simulation::Output) owns them between writes. */
int gdf_snapshot_counters(gdf_ctx* ctx, const int64_t* set2, const double* set2d,
                          int64_t* get2, double* get2d);
/* the column writer on host arrays (no device call): nhydro gas lines then nstar star lines */
int gdf_format_column(const char* path, int nhydro, int nstar, double t, const double* r,
                      const double* v, const double* m, const double* h, const double* rho,
                      const double* u);
/* the Seren unformatted writer on host arrays (Simulation::WriteSerenUnformSnapshotFile,
 * src/Common/SimulationIO.hpp:2009-2249, serial build, FLOAT = double, dimensionless units):
 * counts3 = {Noutsnap, Nsteps, Noutlitesnap}, times4 = {t, tsnaplast, mmean, tlitesnaplast};
 * porig may be NULL (ids 0..nhydro-1); stars follow the gas in r, v, m, h */
int gdf_format_su(const char* path, int nhydro, int nstar, const int64_t* counts3,
                  const double* times4, double h_fac, const int* porig, const double* r,
                  const double* v, const double* m, const double* h, const double* rho,
                  const double* u);
/* the Seren formatted writer on host arrays (Simulation::WriteSerenFormSnapshotFile,
 * src/Common/SimulationIO.hpp:993-1222), arguments as gdf_format_su */
int gdf_format_sf(const char* path, int nhydro, int nstar, const int64_t* counts3,
                  const double* times4, double h_fac, const int* porig, const double* r,
                  const double* v, const double* m, const double* h, const double* rho,
                  const double* u);
/* the Seren lite writer on host arrays (Simulation::WriteSerenLiteSnapshotFile,
 * src/Common/SimulationIO.hpp:2264-2497: single-precision r, m, h, rho, u and star records):
 * counts2 = {Noutsnap, Nsteps}, times3 = {t, tsnaplast, mmean} */
int gdf_format_slite(const char* path, int nhydro, int nstar, const int64_t* counts2,
                     const double* times3, double h_fac, const double* r, const double* v,
                     const double* m, const double* h, const double* rho, const double* u);
/* simulation time, current timestep and step count */
int gdf_time(const gdf_ctx* ctx, double* t, double* timestep, int64_t* nsteps);

/* ---- parity taps (no reference counterpart: the reference's own debug verification,
 * NeighbourManager::VerifyNeighbourList, NeighbourManager.h:574-684, checks the same sets) */
/* exact neighbour sets after the last force pass, ascending ids per particle, import order.
 * mode 0 = gather {j : r2 < k^2 h_i^2}; mode 1 = scatter-gather (NeighbourManager.h:521), by a
 * fresh walk of the tree, for the particles flagged active in a block-timestep run (every
 * particle otherwise; the other rows are empty); mode 2 (mixed path) = the scatter-gather sets
 * the last force pass itself evaluated, for that pass's targets, decoded from the candidate
 * lists it read with its FP32 tests.  offsets has N+1 entries.  With ids == NULL returns the total entry count; otherwise fills ids
 * and returns the total, or GDF_ERR_CAPACITY if cap is smaller. */
int64_t gdf_neighbour_sets(gdf_ctx* ctx, int mode, int64_t* offsets, int32_t* ids,
                           int64_t cap);
/* tree arrays in device (sorted) order; any pointer may be NULL.  perm: sorted slot ->
 * import index [N]; keys [N]; left/right/first/last [N-1] internal nodes (children in the
 * unified index: internal n, leaf l -> N-1+l); mass [2N-1], com [3(2N-1)], quad [5(2N-1)],
 * bbmin/bbmax [3(2N-1)], hmax [2N-1]; group first/count/node [ngroups] */
int gdf_tree_arrays(gdf_ctx* ctx, int32_t* perm, uint64_t* keys, int32_t* left,
                    int32_t* right, int32_t* first, int32_t* last, double* mass, double* com,
                    double* quad, double* bbmin, double* bbmax, double* hmax,
                    int32_t* group_first, int32_t* group_count, int32_t* group_node,
                    int* ngroups);
/* interaction counts of the last passes, for the roofline: {gather pairs of the converged
 * density evaluation, scatter-gather pairs, particle-particle gravity terms,
 * particle-cell gravity terms, density evaluations incl. iterations, candidate-target
 * distance tests of the mixed hydro kernel (it decides the scatter-gather sets itself)} */
int gdf_interaction_counts(gdf_ctx* ctx, int64_t* out6);
/* per-group gravity interaction counts of the last gravity pass [ngroups] each */
int gdf_gravity_group_counts(gdf_ctx* ctx, int32_t* ncell, int32_t* nnear);

/* star-star direct sums for every particle of the context, writing a, adot and gpot:
 * smoothed = 0 replaces Nbody::CalculateDirectGravForces (src/Nbody/Nbody.cpp:233-293),
 * smoothed = 1 NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces
 * (src/Nbody/NbodyLeapfrogKDK.cpp:78-150) */
int gdf_nbody_direct_forces(gdf_ctx* ctx, int smoothed);

/* ---- stars beside the gas: SphSimulation with an N-body integrator (Nbody<3>::stardata,
 * src/Headers/Nbody.h; the N-body half of SphSimulation::MainLoop, SphSimulation.cpp:595,
 * 775-833, 881).  The stars ride in the gas context, in import order, in FP64.
 * gdf_setup_simulation and gdf_main_loop then also give the gas the stars' pull (GradhSph::
 * ComputeStarGravForces, GradhSph.cpp:699-743) and the stars' zeta terms, the stars the gas's
 * pull (HydroTree::UpdateAllStarGasForces, HydroTree.cpp:552-654, every gas particle summed
 * directly: the reference's brute-force search) and each other's, and advance the stars by
 * NbodyLeapfrogKDK (src/Nbody/NbodyLeapfrogKDK.cpp:253-401), with one timestep level (Nlevels
 * = 1); the per-method entry points below also serve a caller's block-timestep loop. ---- */
/* Simulation::ImportArray's star half / the star lines of a column snapshot: nstar stars,
 * r, v (nstar,3) row-major, m, h nstar doubles; replaces earlier stars */
int gdf_set_stars(gdf_ctx* ctx, int nstar, const double* r, const double* v, const double* m,
                  const double* h);
int gdf_num_stars(const gdf_ctx* ctx);
/* a star field in import order: r v a adot r0 v0 a0 (nstar,3), m h gpot (nstar) */
int gdf_get_star_field(gdf_ctx* ctx, const char* name, double* out);
/* HydroTree::UpdateAllStarGasForces: adds the gas's pull on every star to a and gpot */
int gdf_update_all_star_gas_forces(gdf_ctx* ctx);
/* the stars' zeroing and star-star sums of the MainLoop's N-body iteration (SphSimulation.cpp:
 * 778-807): zero_first clears a, adot, gpot (before gdf_update_all_star_gas_forces); else the
 * star-star sums (Nbody::CalculateDirectGravForces or NbodyLeapfrogKDK::
 * CalculateDirectSmoothedGravForces, per nbody_softening) are added */
int gdf_nbody_star_forces(gdf_ctx* ctx, int zero_first);
/* NbodyLeapfrogKDK::AdvanceParticles, CorrectionTerms and EndTimestep(t) for the stars
 * (NbodyLeapfrogKDK.cpp:253-370): a step runs from the stars' last end-step time (their
 * import time before the first) to t */
int gdf_nbody_advance_particles(gdf_ctx* ctx, double t);
int gdf_nbody_correction_terms(gdf_ctx* ctx, double t);
int gdf_nbody_end_timestep(gdf_ctx* ctx, double t);

/* ---- serving pool: no counterpart in the reference, which runs one simulation per process.
 * Independent requests of one configuration class (one gdf_config) share a device context:
 * each is set up alone, then joins a batch whose steps advance every resident request by one
 * of its own global timesteps with one launch of each pass over all of them.  max_particles
 * bounds the resident particles, the admission limit on HBM footprint (about 1.7 KB of device
 * state per particle of capacity, plus 128 MiB of walk stacks per pool).  Requests need
 * Nlevels = 1 and no direct N-body sums. */
typedef struct gdf_pool gdf_pool;
int gdf_pool_create(const gdf_config* cfg, int64_t max_particles, gdf_pool** out);
void gdf_pool_destroy(gdf_pool* pool);
const char* gdf_pool_last_error(const gdf_pool* pool);
/* admission: GDF_ERR_CAPACITY (nothing changed) when the request does not fit beside the
 * resident ones; the caller keeps it queued.  *id receives the request id. */
int gdf_pool_admit(gdf_pool* pool, int N, const double* r, const double* v, const double* m,
                   const double* u, int nsteps, int* id);
/* one step of every resident request; requests that complete their steps leave the batch with
 * their final fields kept.  out2 = {requests stepped, particle-steps} */
int gdf_pool_step(gdf_pool* pool, int64_t* out2);
/* out3 = {0 unknown, 1 resident, 2 finished; steps done; particles}; *t its simulation time */
int gdf_pool_status(const gdf_pool* pool, int id, int64_t* out3, double* t);
/* a per-particle field of a resident request (any gdf_get_field name) or of a finished one
 * (r v a atree m h rho u dudt div_v gpot), in the request's import order */
int gdf_pool_get_field(gdf_pool* pool, int id, const char* name, double* out);
/* forget a finished request's fields */
int gdf_pool_release(gdf_pool* pool, int id);
int64_t gdf_pool_resident_particles(const gdf_pool* pool);

/* Simulation::CalculateDiagnostics as gdf_calculate_diagnostics, on arrays the caller owns in
 * device memory, allocating nothing: no counterpart in the reference; for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  r, v: (N,3) row-major; m, u, gpot: N doubles; out: 11 doubles of device memory
 * receiving {Etot, ketot, utot, gpetot, mom[3], angmom[3], mtot}; utot enters Etot when
 * hydro_forces, gpetot when self_gravity.  All arrays must be device-accessible (device,
 * pinned host or managed memory), 1 <= N <= 2^30.  Asynchronous on stream (a hipStream_t, or
 * NULL for the null stream).  The block partials live in one fixed 88 KiB device array per GPU,
 * shared by every call on that GPU: calls that overlap on one GPU race.  On failure the
 * message is gdf_last_error(NULL). */
int gdf_calculate_diagnostics_noalloc(int N, const double* r, const double* v, const double* m,
                                      const double* u, const double* gpot, int hydro_forces,
                                      int self_gravity, double* out, void* stream);

/* event timing of the named passes of the last step, milliseconds:
 * {tree, density, hydro, gravity, integrate, gravity walk kernel alone (summed over its
 * launches), number of gravity walk launches in the step (block timesteps repeat the force pass
 * for particles whose step CheckTimesteps ends early), hydro pair kernel alone (summed), number
 * of hydro pair-kernel launches} */
int gdf_pass_times(gdf_ctx* ctx, double* out9);
int gdf_set_timing(gdf_ctx* ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif
