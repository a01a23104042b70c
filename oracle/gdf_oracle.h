/* CPU oracle for the GANDALF particle hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C (OpenMP) restatement of the reference algorithms, used as the checker for the
 * HIP path and as the CPU baseline timed beside it.  Only tests/, the smoke check and the
 * benchmark's cpu_baseline leg may load it.  The product path never calls into it.
 *
 * Every function cites the reference file:line it restates.  Layout conventions: positions,
 * velocities and accelerations are (N,3) row-major double arrays (r[3*i+k]); everything else
 * is a length-N double array.  Particles are identified by their index in the input arrays.
 */
#ifndef GDF_ORACLE_H
#define GDF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int kernel;              /* 0 = M4, 1 = quintic */
  int tabulated;           /* 1 = 1000-entry nearest-lower tables (SmoothingKernel.h:548-756) */
  double h_fac, h_converge;
  int eos;                 /* 0 = energy_eqn (adiabatic), 1 = isothermal, 2 = barotropic */
  double gamma, temp0, mu_bar, rho_bary;
  int energy_integration;  /* 1 when u is integrated (energy_eqn) */
  double alpha_visc, beta_visc;
  int hydro_forces, self_gravity;
  double thetamaxsqd;
  int nleafmax;            /* leaf bucket size of the tree walks */
  int periodic;            /* 1 = all three axes periodic */
  double boxmin[3], boxmax[3];
  double courant_mult, accel_mult, energy_mult;
  int nlevels;             /* block timestep levels (1 = global timestep) */
  int level_diff_max;      /* max. level difference to neighbours (Simulation.cpp:1209) */
  int grav_direct;         /* 1: the sim driver's gravity is the exact direct sum, as the
                              reference's neib_search = bruteforce (BruteForceTree.cpp:214) */
  int multipole;           /* cell terms of the gravity walk (Parameters.cpp:309): 0 quadrupole,
                              1 monopole, 2 fast_monopole, 3 fast_quadrupole */
  int gravity_mac;         /* opening criterion (Parameters.cpp:308): 0 geometric, 1 gadget2,
                              2 eigenmac */
  double macerror;         /* their error parameter (Parameters.cpp:314) */
} orc_params;

enum { ORC_W0 = 0, ORC_W1, ORC_WOMEGA, ORC_WZETA, ORC_WGRAV, ORC_WPOT,
       ORC_W0_S2, ORC_WOMEGA_S2, ORC_WZETA_S2 };

/* kernel functions (SmoothingKernel.h:132-408, tabulated :583-637) */
double orc_kernel_eval(const orc_params* p, int which, double s);
double orc_kernel_range(const orc_params* p);

/* initial smoothing length guess (Sph.cpp:76-113, 3D branch) */
double orc_h_guess(const orc_params* p, int N, const double* r);

/* exact neighbour sets, ascending j per particle.  mode 0 = gather {j : r2 < k^2 h_i^2};
 * mode 1 = scatter-gather {j : r2 < k^2 h_i^2 or r2 < k^2 h_j^2} (NeighbourManager.h:521);
 * mode 2 = search {j : r2 <= k^2 h_i^2} with h the search length.  Returns total entries,
 * or -(needed) if cap is too small.  offsets has N+1 entries. */
int64_t orc_neighbour_sets(const orc_params* p, int N, const double* r, const double* h,
                           int mode, int64_t* offsets, int32_t* ids, int64_t cap);

/* grad-h h-rho iteration for every particle (GradhSph.cpp:142-326 driven as
 * GradhSphTree.cpp:83-271), then the EOS (GradhSph.cpp:335-347).  h is in/out; u is in/out
 * (set by isothermal/barotropic EOS).  nopbc=1 ignores periodicity (first setup pass).
 * Returns 0, or -1 if an h iteration failed to converge. */
int orc_compute_h(const orc_params* p, int N, const double* r, const double* m, int nopbc,
                  double* h, double* rho, double* invomega, double* zeta, double* hfactor,
                  double* u, double* sound, double* pressure);

/* hydro forces over the scatter-gather sets; variant 0 = ComputeSphHydroForces
 * (GradhSph.cpp:361-460), variant 1 = hydro part of ComputeSphHydroGravForces (:474-585).
 * Overwrites a, dudt, div_v. */
void orc_hydro_forces(const orc_params* p, int variant, int N, const double* r,
                      const double* v, const double* m, const double* h, const double* rho,
                      const double* invomega, const double* pressure, const double* sound,
                      const double* hfactor, double* a, double* dudt, double* div_v);

/* FP64 direct O(N^2) kernel-softened gravity, the theta -> 0 limit of the walk: neighbour
 * pairs as GradhSph.cpp:555-565, others as :657-690, gpot seeded with the self term as
 * GradhSphTree.cpp:512.  Overwrites atree and gpot. */
void orc_direct_gravity(const orc_params* p, int N, const double* r, const double* m,
                        const double* h, const double* zeta, const double* hfactor,
                        double* atree, double* gpot);

/* Plummer-softened point-mass direct sum (G=1): a_i = sum m_j dr/(r^2+eps^2)^1.5 */
void orc_direct_plummer(int N, const double* r, const double* m, double eps,
                        const int32_t* targets, int ntargets, double* a, double* pot);

/* Star-star direct sums of the N-body integrator, every star active, outputs overwritten:
 * Nbody::CalculateDirectGravForces (src/Nbody/Nbody.cpp:233-293), unsoftened with jerk, and
 * NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces (src/Nbody/NbodyLeapfrogKDK.cpp:78-150),
 * kernel-softened over the mean smoothing length 2/(h_i + h_j). */
void orc_nbody_direct(int N, const double* r, const double* v, const double* m, double* a,
                      double* adot, double* gpot);
void orc_nbody_direct_smoothed(const orc_params* p, int N, const double* r, const double* v,
                               const double* m, const double* h, double* a, double* adot,
                               double* gpot);

/* ---- Ewald periodic gravity (fully periodic box) ----
 * Ewald::Ewald's correction field (Ewald.cpp:42-420, ewald_periodicity = 7): field holds
 * 4*Ng[0]*Ng[1]*Ng[2] doubles, component-major (potential, a_x, a_y, a_z), or is NULL to get
 * Ng and dI only */
int orc_ewald_table(const double L[3], int nseries, int ngrid, double ewald_mult,
                    double efratio, double* field, int Ng[3], double dI[3]);
typedef struct {
  int Ng[3];
  double dI[3];
  double L[3];
  const double* field;
} orc_ewald;
/* Ewald::CalculatePeriodicCorrection (Ewald.cpp:500-650): the correction of mass m at
 * dr = r_source - r_target */
void orc_ewald_corr(const orc_ewald* e, double m, const double* dr, double* acorr,
                    double* gpotcorr);
/* exact periodic gravity: every pair's nearest-image term plus its Ewald correction */
void orc_direct_gravity_periodic(const orc_params* p, const orc_ewald* e, int N,
                                 const double* r, const double* m, const double* h,
                                 const double* zeta, const double* hfactor, double* atree,
                                 double* gpot);

/* ---- stars beside the gas (Nbody<3>::stardata of an SphSimulation) ----
 * r (n,3), m, h n doubles; softening = nbody_softening, conservative =
 * conservative_sph_star_gravity */
typedef struct {
  int n;
  const double *r, *m, *h;
  int softening, conservative;
} orc_stars;
/* orc_compute_h with the stars' zeta terms (GradhSph.cpp:282-307) */
int orc_compute_h_stars(const orc_params* p, int N, const double* r, const double* m,
                        int nopbc, const orc_stars* st, double* h, double* rho,
                        double* invomega, double* zeta, double* hfactor, double* u,
                        double* sound, double* pressure);
/* GradhSph::ComputeStarGravForces for every gas particle: adds to a and gpot */
void orc_gas_from_stars(const orc_params* p, int N, const double* r, const double* h,
                        const orc_stars* st, double* a, double* gpot);
/* HydroTree::UpdateAllStarGasForces through the brute-force search: writes the stars' a, gpot */
void orc_stars_from_gas(const orc_params* p, const orc_stars* st, int N, const double* r,
                        const double* m, const double* h, double* a, double* gpot);

/* ---- CPU twin of the device tree (Morton keys, LBVH, bottom-up moments, groups) ----
 * Arrays are sized: perm[N], parent[2N-1], left/right/first/last[N-1] (internal nodes),
 * nodes' moments [2N-1] in the unified index (internal 0..N-2, leaf l -> N-1+l).
 * group_first/group_count/group_node sized N (returns number of groups). */
typedef struct {
  int N;
  int32_t *perm;                         /* sorted slot -> input index */
  uint64_t *keys;                        /* sorted keys */
  int32_t *left, *right, *first, *last;  /* internal nodes */
  int32_t *parent;                       /* unified index */
  double *mass, *com, *quad, *bbmin, *bbmax, *hmax, *cdistsqd;  /* com/bb x3, quad x5 */
  int32_t *group_first, *group_count, *group_node;
  int ngroups;
} orc_tree;

int orc_tree_alloc(orc_tree* t, int N);
void orc_tree_free(orc_tree* t);
/* r, m, h in input order; builds everything.  group_max = particles per walk group. */
int orc_tree_build(const orc_params* p, orc_tree* t, const double* r, const double* m,
                   const double* h, int group_max);
/* refresh hmax bottom-up after h changed (KDTree.cpp:1128-1208 analogue); h in input order */
void orc_tree_refit_h(const orc_params* p, orc_tree* t, const double* h);

/* tree gravity with the device walk's acceptance rules; r,m,h,zeta,hfactor in input order.
 * If ncell/nnear are non-null they receive per-group interaction counts. */
/* atree_prev (N x 3) and gpot_prev (N): the particles' tree acceleration and potential when the
 * tree was stocked, for the gadget2 and eigenmac criteria (NULL for the geometric one) */
void orc_tree_gravity(const orc_params* p, const orc_tree* t, const double* r,
                      const double* m, const double* h, const double* zeta,
                      const double* hfactor, double* atree, double* gpot,
                      int32_t* ncell, int32_t* nnear, const double* atree_prev,
                      const double* gpot_prev);

/* one quadrupole cell term (NeighbourSearch.h:423-475): a cell of mass mc, centre of mass com
 * and traceless moments q[5] acting on a point rp; adds to agrav[3] and *gpot */
void orc_cell_quad_force(const double* com, const double* q, double mc, const double* rp,
                         double* agrav, double* gpot);
/* Tree::open_cell_for_gravity (Tree.h:413-432) past the geometric test, for p->gravity_mac
 * gadget2 (1) or eigenmac (2): 1 if a cell (box half-widths half, mass mc, moments q[5]) at
 * squared distance dmac2 is opened, given the walking group's factor gfac (smallest |atree| or
 * largest gpot^(-2/3)) */
int orc_cell_mac_opens(const orc_params* p, const double* half, double mc, const double* q,
                       double dmac2, double gfac);
/* one monopole cell term (ComputeCellMonopoleForces, NeighbourSearch.h:350-377) */
void orc_cell_mono_force(const double* com, double mc, const double* rp, double* agrav,
                         double* gpot);
/* the fast multipole expansion (FastMultipoleForces, NeighbourSearch.h:478-760, first order) of
 * n cells at dcom (n x 3, relative to the expansion point), moments q (n x 5), masses m; quad = 0
 * monopole parts only, 1 with the quadrupole parts.  coef[10] = (pot, a[3], tensor xx xy yy xz
 * yz zz); agrav (np x 3) and gpot (np) = the series at the points d (np x 3, relative) */
void orc_fast_multipole(int n, const double* dcom, const double* q, const double* m, int quad,
                        int np, const double* d, double* coef, double* agrav, double* gpot);

/* ---- integrator (SphLeapfrogKDK.cpp:76-272) and timesteps (SphIntegration.cpp:81-134) ---- */
void orc_advance(int N, double dt, int energy_integration, const double* r0,
                 const double* v0, const double* a0, const double* u0, const double* dudt0,
                 double* r, double* v, double* u);
void orc_end_step(int N, double dtstep, int energy_integration, const double* a,
                  const double* dudt, double* v, double* u, double* r0, double* v0,
                  double* a0, double* u0, double* dudt0, const double* r);
double orc_timesteps(const orc_params* p, int N, const double* h, const double* sound,
                     const double* div_v, const double* a, const double* u,
                     const double* dudt, double* dt_out);
void orc_wrap_periodic(const orc_params* p, int N, double* r);

/* diagnostics (SimAnalysis.hpp:52-185): out = {Etot, ketot, utot, gpetot, mom[3],
 * angmom[3], mtot} */
void orc_diagnostics(const orc_params* p, int N, const double* r, const double* v,
                     const double* m, const double* u, const double* gpot, double* out);

/* ---- full step driver (SphSimulation.cpp:204-880 with global timesteps) ---- */
typedef struct orc_sim orc_sim;
orc_sim* orc_sim_create(const orc_params* p, int N, const double* r, const double* v,
                        const double* m, const double* u);
void orc_sim_destroy(orc_sim* s);
/* stars beside the gas, leapfrog KDK (NbodyLeapfrogKDK), before orc_sim_setup; orc_sim_get
 * then also answers star_r star_v star_a star_adot star_gpot */
int orc_sim_set_stars(orc_sim* s, int n, const double* r, const double* v, const double* m,
                      const double* h, int softening, int conservative, double nbody_mult);
int orc_sim_setup(orc_sim* s);
int orc_sim_step(orc_sim* s);
/* field names: r v a atree m h rho u dudt div_v gpot invomega zeta sound pressure hfactor */
int orc_sim_get(orc_sim* s, const char* field, double* out);
double orc_sim_time(orc_sim* s);
double orc_sim_timestep(orc_sim* s);
/* block-timestep state: iout = n, level_max, level_step, nresync; dout = dt_max, number of
 * particles that ended a step in the last MainLoop */
void orc_sim_block_state(orc_sim* s, int* iout, double* dout);

#ifdef __cplusplus
}
#endif
#endif
