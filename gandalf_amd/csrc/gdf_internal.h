// Internal declarations shared by the HIP translation units of libgandalf_gpu.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <string>
#include <vector>

#include "../../include/gandalf_gpu.h"

namespace gdf {

// Walk groups hold at most one wavefront of targets.
constexpr int kWave = 64;
constexpr int kMaxGroup = 64;
// Persistent wavefronts of the candidate walks (3 per SIMD: 4, 6 and 8 were slower, more walks
// in flight contending for the caches) and of the gravity walk (4 per SIMD); walk kernels grab
// groups from a counter.
constexpr int kWalkSlots = 3072;
constexpr int kGravSlots = 4096;
// Per-slot stack capacity of the tree walks (node indices); the stack regions serve both walks.
constexpr int kStackCap = 8192;
constexpr int kStackSlots = kWalkSlots > kGravSlots ? kWalkSlots : kGravSlots;

// Per-particle persistent fields, permuted with the particles at every tree build.
// F_TLAST, F_PDT, F_DTNEXT: start time, length and next length of the particle's block step
// (Particle.h:166-167, dt).
enum PField {
  F_X, F_Y, F_Z, F_VX, F_VY, F_VZ, F_X0, F_Y0, F_Z0, F_VX0, F_VY0, F_VZ0,
  F_AX0, F_AY0, F_AZ0, F_M, F_H, F_U, F_U0, F_DUDT0, F_TLAST, F_PDT, F_DTNEXT, F_NPERSIST
};
// Per-particle integer state, permuted with the particles: block-timestep level, neighbour
// level, integer step length and start (Particle.h:140-143), and the active / end-of-step
// flags (Flags.h active, end_timestep) as 0/1.
enum IField { I_LEVEL, I_LEVELNEIB, I_NSTEP, I_NLAST, I_ACT, I_END, I_SEG, I_NINT };
// Serving batches: up to kMaxSeg independent simulations share one context, each a contiguous
// particle segment; the segment id prefixes the Morton key (7 bits + 19 bits per axis), so the
// sort keeps segments apart and each segment is one subtree of the LBVH.
constexpr int kMaxSeg = 128;

// Derived per-particle fields, recomputed every step after the permutation.
enum DField {
  D_AX, D_AY, D_AZ, D_ATX, D_ATY, D_ATZ, D_RHO, D_DUDT, D_DIVV, D_GPOT, D_INVOMEGA,
  D_ZETA, D_SOUND, D_PRESSURE, D_HFACTOR, D_HSEARCH, D_DT, D_ADOTX, D_ADOTY, D_ADOTZ,
  // the gravity walk's own sums when it runs beside the hydro pass: acceleration, potential,
  // and the self term (m/h) wpot(0)
  D_GWX, D_GWY, D_GWZ, D_GWP, D_GWP0,
  D_NDERIVED
};

// Stars of a gas + star run (NbodyParticle, src/Headers/NbodyParticle.h), in import order, one
// device array per field: position, velocity, acceleration, jerk, their values at the start of
// the step, mass, smoothing length, potential.
enum SField {
  S_X, S_Y, S_Z, S_VX, S_VY, S_VZ, S_AX, S_AY, S_AZ, S_ADX, S_ADY, S_ADZ,
  S_X0, S_Y0, S_Z0, S_VX0, S_VY0, S_VZ0, S_AX0, S_AY0, S_AZ0, S_M, S_H, S_GPOT, S_N
};
// the star-from-gas sums take chunks of this many gas particles per workgroup
constexpr int kStarChunk = 2048;

// Smoothing-kernel description passed by value to device code.
struct KernelDesc {
  int type;            // 0 M4, 1 quintic
  int tabulated;
  double range, rangesqd, norm, resinv, resinvsqd;
  int res;
  const double* tab;   // 9*res doubles (device), ORC order: w0 w1 womega wzeta wgrav wpot w0_s2 womega_s2 wzeta_s2
  const float* tabf;   // float copy for the mixed path
};

struct PhysParams {
  KernelDesc kern;
  double h_fac, h_converge, gamma, gammam1, temp0, mu_bar, invrho_bary;
  int eos, energy_integration;
  double alpha_visc, beta_visc;
  int periodic;
  double boxsize[3], boxhalf[3], boxmin[3], boxmax[3];
  double thetamaxsqd;
  int nleaf;
  double courant_mult, accel_mult, energy_mult;
  int hydro_forces, self_gravity;
  int nbody;                 // pure N-body: stars only, no gas pass
  double eps2;               // Plummer softening length squared of the N-body tree walk
  double nbody_mult;
  int hrho_itmax;            // fixed-point iterations before bisection (reference: 30)
  int multipole;             // GDF_MULTIPOLE_*: the gravity walk's accepted-cell terms
  int gravity_mac;           // GDF_MAC_*: its opening criterion
  double macerror;
  // Ewald periodic gravity (periodic box with self-gravity): correction table of eng[0] x
  // eng[1] x eng[2] nodes, x fastest, each (pot, ax, ay, az) of a unit mass; node spacing 1/edI
  int ewald;
  int eng[3];
  double edI[3];
  const double4* etab; const float4* etabf;
};

// Tree node arrays in the unified index (internal 0..N-2, leaf l -> N-1+l).
struct TreeView {
  int N;
  const int* left; const int* right; const int* first; const int* last;
  const double4* geo_c;   // (cx, cy, cz, hmax)
  const double4* geo_h;   // (hx, hy, hz, cdistsqd)
  const double* mp;       // 10 doubles per node: comx comy comz m q0..q4 pad
};

// Candidate set of the SPH passes, per walk group g: idx[g][b*64 + c] is candidate c of batch b,
// mask[g][b*64 + l] the 64-bit set of batch-b candidates inside target l's range, ncand[g] the
// number of candidates.  capb batches are reserved per group.
struct CandSet {
  int* idx = nullptr;
  unsigned long long* mask = nullptr;
  int* ncand = nullptr;
  int capb = 0;
  int gcap = 0;
};

}  // namespace gdf

struct gdf_ctx {
  gdf_config cfg;
  gdf::PhysParams phys;
  std::string err;
  hipStream_t stream = nullptr;
  int device = 0;

  int N = 0, Ncap = 0;
  double* pf[gdf::F_NPERSIST] = {};      // persistent fields (sorted order)
  double* pf_alt[gdf::F_NPERSIST] = {};  // permutation targets
  double* df[gdf::D_NDERIVED] = {};      // derived fields
  double* df_alt[gdf::D_NDERIVED] = {};  // permutation targets (block timesteps only: inactive
                                         // particles keep their derived state across steps)
  int* pi[gdf::I_NINT] = {};             // integer state (sorted order)
  int* pi_alt[gdf::I_NINT] = {};
  int* iorig = nullptr; int* iorig_alt = nullptr;
  double4* posm = nullptr;               // packed (x, y, z, m) for neighbour gathers
  double4* hrecA = nullptr;              // hydro records: (vx, vy, vz, 1/h)
  double4* hrecB = nullptr;              //                (1/rho, P*invomega/rho^2, sound, hfactor)
  double4* grec = nullptr;               // gravity records: (h, zeta, hfactor, 0)
  int* status = nullptr;

  // tree
  uint64_t* keys = nullptr; uint64_t* keys_alt = nullptr;
  int* vals = nullptr; int* vals_alt = nullptr;
  void* sort_tmp = nullptr; size_t sort_tmp_bytes = 0;
  int* left = nullptr; int* right = nullptr; int* first = nullptr; int* last = nullptr;
  int* parent = nullptr; int* visit = nullptr;
  int4* lnk = nullptr;          // internal node: (first, last, left, right) in one 16-byte record
  int* front = nullptr;                                // refit frontier [N] + count
  double* bbmin = nullptr; double* bbmax = nullptr;     // 3 per node
  double4* geo_c = nullptr; double4* geo_h = nullptr;
  double* mp = nullptr;                                // 10 per node
  int* gstart = nullptr;                               // per-slot group-start marker
  int* grp_first = nullptr; int* grp_count = nullptr; int* grp_node = nullptr;
  int* nsel = nullptr; void* sel_tmp = nullptr; size_t sel_tmp_bytes = 0;
  int ngroups = 0;
  bool tree_valid = false;       // the tree matches the particles' current state
  bool topo_valid = false;       // the tree's topology and groups match the particle order
  bool rebuild_tree = true;      // SphSimulation's rebuild_tree: the next main loop rebuilds

  // per-group candidate sets: gather (h-rho search) and scatter-gather (forces)
  gdf::CandSet cs_g, cs_s;
  bool sg_valid = false;
  bool sg_masks = false;        // the scatter-gather candidate set holds its masks
  // mixed path: the last density pass's candidate lists (cs_g) also hold every scatter-gather
  // pair of the final smoothing lengths (walked with the search radii as bounds); the force
  // pass then reads them instead of walking the tree again (sg_from_union)
  bool sg_union = false;
  bool sg_from_union = false;
  // the last mixed block-step force pass's targets (its active flags, kept for decoding the
  // sets it evaluated); sg_tgt_all: every particle was a target (global timesteps)
  int* sg_tgt = nullptr;
  int sg_tgt_cap = 0;
  bool sg_tgt_all = true;

  // walk scratch
  int* stacks = nullptr;
  int* stacks2 = nullptr;      // the gravity walk's stacks when it runs beside the hydro pass
  // walk groups with at least one target, in ascending order: [0, Ncap] for the gather walk
  // of the h-rho rounds, [Ncap + 1, 2 Ncap + 1] for the active particles of a block step (the
  // scatter-gather walk and the gravity walk); each list's count follows its Ncap entries
  int* glist = nullptr;
  int* counters = nullptr;     // [0..7] flags, [4..11] walk range counters, [12..13] block
                               // reductions; [16..23] flags and [24..31] range counters of the
                               // gravity walk on the second stream
  hipStream_t stream2 = nullptr;             // gravity walk beside the hydro pass
  hipEvent_t ev_dens = nullptr, ev_grav = nullptr;
  int* h_counters = nullptr;   // pinned mirror
  long long* icount = nullptr; // [8] interaction counters
  int32_t* grav_nc = nullptr; int32_t* grav_nn = nullptr;
  double* red = nullptr;       // reduction partials
  double* h_red = nullptr;     // pinned
  double* tmpd = nullptr;      // scratch N*3 doubles for field export

  // kernel tables
  double* ktab = nullptr; float* ktabf = nullptr;

  // simulation driver state (SphSimulation)
  double t = 0.0, timestep = 0.0, tlast = 0.0;
  int64_t nsteps = 0;
  bool setup_done = false;
  // block timesteps (Simulation.h:187-215): integer time n, nresync, level_max, level_step
  int nlevels = 1, level_diff_max = 1;
  int n_int = 0, nresync = 0, level_max = 0, level_step = 0;
  double dt_max = 0.0;
  long long finished_last = 0;   // particles that ended a step in the last MainLoop
  // serving batch (nseg > 0): segment s holds particles [seg_off[s], seg_off[s+1]); per-segment
  // bounds, subtree roots and step lengths live on the device
  int nseg = 0;
  std::vector<int> seg_off;
  int* d_seg_off = nullptr;      // [kMaxSeg + 1]
  int* seg_root = nullptr;       // [kMaxSeg]
  int* grp_root = nullptr;       // [Ncap]: subtree root of each walk group's segment
  // [Ncap]: each walk group's factor of the gadget2 / eigenmac opening criterion, from its
  // particles' atree / gpot at the last tree build or refit
  double* grp_gmac = nullptr;
  double* seg_dt = nullptr;      // [kMaxSeg]: drift / kick length of the segment's step
  double* seg_bnd = nullptr;     // [4 kMaxSeg]: lower corner and scale of its Morton grid

  // timing
  bool timing = false;
  hipEvent_t ev[12] = {};      // pass-boundary events
  double pass_ms[7] = {};       // ... [5] gravity walk kernel, [6] hydro pair kernel
  bool mark_passes = true;     // record the pass-boundary events (first force pass of a step)
  int grav_launches = 0;       // gravity walk launches in the current step
  long long n_active = 0;      // particles the current force pass computes (block steps: active)
  int hydro_launches = 0;      // hydro pair-kernel launches in the current step
  // pair-kernel launches bracketed by event pairs, read back at the end of the step (or when
  // the pairs run out) rather than with a host wait after every launch
  static constexpr int kKevPairs = 16;
  hipEvent_t kev[2*kKevPairs] = {};
  int kev_kind[kKevPairs] = {};
  int nkev = 0;
  long long icounts_last[6] = {};
  // Seren binary snapshot header state (Simulation.h Noutsnap, Noutlitesnap, tsnaplast,
  // tlitesnaplast) and the file's particle ids, carried from a read to the next write
  int64_t snap_nout = 0, snap_nlite = 0;
  double snap_tlast = 0.0, snap_tlitelast = 0.0;
  std::vector<int> snap_porig;

  // captured graphs of the tree build's launches, one per buffer arrangement (the build's
  // swaps alternate between two); graph_failed: a capture failed, the context builds eagerly
  static constexpr int kBuildGraphs = 2;
  hipGraphExec_t bg_exec[kBuildGraphs] = {};
  std::vector<const void*> bg_key[kBuildGraphs];
  int bg_next = 0;
  bool graph_failed = false;
  int graph_why = 0;             // why the last capture was given up (0: it was not)
  long long graph_replays = 0;

  // Ewald correction table (device copies, FP64 and FP32)
  double4* ewald_tab = nullptr;
  float4* ewald_tabf = nullptr;

  // stars beside the gas (gdf_set_stars): star-gas gravity both ways, star-star sums, the
  // stars' leapfrog KDK (SphSimulation::MainLoop's N-body half)
  int Nstar = 0, Nstar_cap = 0;
  double star_tlast = 0.0;         // the stars' last end-step time (NbodyParticle tlast)
  double* sv[gdf::S_N] = {};
  double* star_part = nullptr;     // per-chunk partial sums of the star-from-gas pass
  size_t star_part_cap = 0;
};

namespace gdf {

int fail(gdf_ctx* c, int code, const std::string& msg);
int check_hip(gdf_ctx* c, hipError_t e, const char* what);
#define GDF_HIP(c, call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return gdf::check_hip((c), e_, #call); } while (0)
#define GDF_TRY(call) do { int rc_ = (call); if (rc_ != GDF_OK) return rc_; } while (0)

// tree (gdf_tree.hip)
int tree_alloc(gdf_ctx* c);
int seg_alloc(gdf_ctx* c);
// context helpers (gdf_capi.hip)
int ctx_alloc_particles(gdf_ctx* c, int ncap);
int ctx_get_field_range(gdf_ctx* c, const char* name, int i0, int n, double* out);
int ctx_field_width(gdf_ctx* c, const char* name);
int seg_upload(gdf_ctx* c);
int tree_build(gdf_ctx* c);
int tree_restock(gdf_ctx* c);
void build_graphs_free(gdf_ctx* c);   // the captured tree-build graphs   // refit of the kept topology to the current particle state
int tree_refit_h(gdf_ctx* c);
// kernel timing (ctx->timing): kind 5 gravity walk, 6 hydro pair kernel.  begin returns the
// event pair's slot (-1 with timing off); collect waits for the recorded pairs and adds them
// to pass_ms and the launch counts
int ktime_begin(gdf_ctx* c, hipStream_t s, int kind);
void ktime_end(gdf_ctx* c, hipStream_t s, int slot);
void ktime_collect(gdf_ctx* c);
// compacted list of the walk groups holding a particle with status[i] == 1 (ascending, count at
// list[Ncap]); the optional per-group outputs of every other group are cleared, and so are
// zero[0, nzero) (nzero <= 256: the flags and queue counters of the walk that follows)
int group_list(gdf_ctx* c, const int* status, int* list, int* ncand, int* gnc, int* gnn,
               int* zero = nullptr, int nzero = 0);
TreeView tree_view(const gdf_ctx* c);

// SPH passes (gdf_sph.hip)
int sph_alloc_lists(gdf_ctx* c);
void sph_free_lists(gdf_ctx* c);
int sph_copy_sg(gdf_ctx* c, std::vector<int>& idx, std::vector<unsigned long long>& mask,
                std::vector<int>& ncand, int& capb);
int sph_density(gdf_ctx* c, int nopbc);
// masks: also record each target's candidate masks (the mixed path's hydro kernel otherwise
// tests its candidates itself)
int sph_build_sg(gdf_ctx* c, bool masks = false);
int sph_hydro(gdf_ctx* c, int variant);

// gravity (gdf_gravity.hip)
int gravity_walk(gdf_ctx* c);
// hydro + self-gravity: the walk on the second stream (after the density pass), then joined:
// atree += walk, gpot = (self + gpot) + walk, a += atree
int gravity_launch(gdf_ctx* c);
int gravity_join(gdf_ctx* c);

// integration (gdf_integrate.hip)
int zero_icount(gdf_ctx* c);    // clears the eight interaction counters of a step
int integ_advance(gdf_ctx* c, double dt);
// block timesteps (gdf_integrate.hip)
int block_init(gdf_ctx* c);
int block_advance(gdf_ctx* c);
int block_zero_active(gdf_ctx* c);
int block_eos_all(gdf_ctx* c);
int block_check_timesteps(gdf_ctx* c, int* activecount);
int block_timesteps(gdf_ctx* c);
int block_end_step(gdf_ctx* c);
int integ_wrap(gdf_ctx* c);
int integ_end_step(gdf_ctx* c, double dtstep);
int integ_timestep(gdf_ctx* c, double* dtmin);
int integ_diagnostics(gdf_ctx* c, double* out);
int integ_max_h(gdf_ctx* c, double* hmax);
int integ_zero_forces(gdf_ctx* c);
// Sph::ZeroAccelerations for the context's timestep mode (every particle, or the active ones
// of a block step plus the thermal state of all)
int integ_zero_accelerations(gdf_ctx* c);
// p[0, n) = v on the context's stream
void fill_int(gdf_ctx* c, int* p, int n, int v);
int integ_add_atree(gdf_ctx* c);
int integ_pack_posm(gdf_ctx* c);
// serving batches: per-segment timestep minima into out[nseg]; drift/kick with seg_dt
int integ_seg_timesteps(gdf_ctx* c, double* out);
int integ_seg_advance(gdf_ctx* c);
int integ_seg_end_step(gdf_ctx* c);
int nbody_direct(gdf_ctx* c, int smoothed);
// stars beside the gas (gdf_stars.hip)
int nbody_direct_stars(gdf_ctx* c, int smoothed);   // star-star sums added to the stars
void stars_free(gdf_ctx* c);
// HydroTree::UpdateAllStarGasForces when with_gas, added to the stars' a and gpot (add) or
// written over a, gpot with adot cleared (!add: the MainLoop's zeroing of the stars)
int stars_gas_on_stars(gdf_ctx* c, int with_gas, int add);
int stars_on_gas(gdf_ctx* c);         // GradhSph::ComputeStarGravForces for the active gas
int stars_advance(gdf_ctx* c, double t);
int stars_correction(gdf_ctx* c, double t);
int stars_end_step(gdf_ctx* c, double t);
int stars_timestep(gdf_ctx* c, double* dtmin);
int stars_diagnostics(gdf_ctx* c, double* acc11);   // adds the stars' sums to acc

// Ewald::Ewald's correction field of a fully periodic box (host/ewald_table.cpp): 4 doubles
// (pot, ax, ay, az) per node, Ng nodes per axis, dI nodes per unit length
int ewald_build_table(const double L[3], int nseries, int ngrid, double ewald_mult,
                      double efratio, std::vector<double>& field, int Ng[3], double dI[3]);

inline int nblocks(long long n, int bs) { return (int) ((n + bs - 1) / bs); }

// gdf_calculate_diagnostics_noalloc: largest particle count, and the message of its last
// failure on this thread (NULL after a successful call)
constexpr long long kNoallocMaxN = 1LL << 30;
const char* noalloc_last_error();

}  // namespace gdf
