// C ABI of libgandalf_gpu: context lifetime, particle import/export and the hot-path entry
// points declared in include/gandalf_gpu.h.  No C++ exception crosses the boundary; errors are
// status codes plus gdf_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {

int fail(gdf_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int check_hip(gdf_ctx* c, hipError_t e, const char* what) {
  if (c) c->err = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
  return e == hipErrorOutOfMemory ? GDF_ERR_NOMEM : GDF_ERR_HIP;
}

namespace {

constexpr int BS = 256;

__global__ void k_scatter_out(int N, int w, const double* f0, const double* f1, const double* f2,
                              const int* __restrict__ iorig, double* __restrict__ out) {
  const int s = blockIdx.x*BS + threadIdx.x;
  if (s >= N) return;
  const int o = iorig[s];
  out[(size_t) o*w] = f0[s];
  if (w == 3) { out[(size_t) o*3 + 1] = f1[s]; out[(size_t) o*3 + 2] = f2[s]; }
}

__global__ void k_gather_in(int N, int w, double* f0, double* f1, double* f2,
                            const int* __restrict__ iorig, const double* __restrict__ in) {
  const int s = blockIdx.x*BS + threadIdx.x;
  if (s >= N) return;
  const int o = iorig[s];
  f0[s] = in[(size_t) o*w];
  if (w == 3) { f1[s] = in[(size_t) o*3 + 1]; f2[s] = in[(size_t) o*3 + 2]; }
}

__global__ void k_iota(int N, int* p) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < N) p[i] = i;
}

// integer per-particle state exported as doubles
int int_field(const char* name) {
  static const char* names[] = {"level", "levelneib", "nstep", "nlast", "active"};
  static const int ids[] = {I_LEVEL, I_LEVELNEIB, I_NSTEP, I_NLAST, I_ACT};
  for (int k = 0; k < 5; k++) if (!strcmp(name, names[k])) return ids[k];
  return -1;
}

__global__ void k_scatter_out_int(int N, const int* __restrict__ src, const int* __restrict__ io,
                                  double* __restrict__ out) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < N) out[io[i]] = (double) src[i];
}

__global__ void k_gather_in_int(int N, int* __restrict__ dst, const int* __restrict__ io,
                                const double* __restrict__ in) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < N) dst[i] = (int) in[io[i]];
}

// resolve a field name to up to three device arrays
bool field_ptrs(gdf_ctx* c, const char* name, double** p, int* w) {
  struct E { const char* n; int src; int f; int w; };
  static const E tab[] = {
    {"r", 0, F_X, 3}, {"v", 0, F_VX, 3}, {"r0", 0, F_X0, 3}, {"v0", 0, F_VX0, 3},
    {"a0", 0, F_AX0, 3}, {"m", 0, F_M, 1}, {"h", 0, F_H, 1}, {"u", 0, F_U, 1},
    {"u0", 0, F_U0, 1}, {"dudt0", 0, F_DUDT0, 1},
    {"a", 1, D_AX, 3}, {"atree", 1, D_ATX, 3}, {"rho", 1, D_RHO, 1}, {"dudt", 1, D_DUDT, 1},
    {"div_v", 1, D_DIVV, 1}, {"gpot", 1, D_GPOT, 1}, {"invomega", 1, D_INVOMEGA, 1},
    {"zeta", 1, D_ZETA, 1}, {"sound", 1, D_SOUND, 1}, {"pressure", 1, D_PRESSURE, 1},
    {"hfactor", 1, D_HFACTOR, 1}, {"dt", 1, D_DT, 1}, {"adot", 1, D_ADOTX, 3},
    {"tlast", 0, F_TLAST, 1}, {"dtstep", 0, F_PDT, 1}};
  for (const E& e : tab) {
    if (strcmp(e.n, name)) continue;
    for (int k = 0; k < e.w; k++) p[k] = e.src == 0 ? c->pf[e.f + k] : c->df[e.f + k];
    *w = e.w;
    return true;
  }
  return false;
}

void free_particles(gdf_ctx* c) {
  build_graphs_free(c);
  for (int f = 0; f < F_NPERSIST; f++) { hipFree(c->pf[f]); hipFree(c->pf_alt[f]); c->pf[f] = c->pf_alt[f] = nullptr; }
  for (int f = 0; f < D_NDERIVED; f++) {
    hipFree(c->df[f]); hipFree(c->df_alt[f]); c->df[f] = c->df_alt[f] = nullptr;
  }
  for (int f = 0; f < I_NINT; f++) { hipFree(c->pi[f]); hipFree(c->pi_alt[f]); c->pi[f] = c->pi_alt[f] = nullptr; }
  void* ptrs[] = {c->iorig, c->iorig_alt, c->posm, c->hrecA, c->hrecB, c->grec, c->status,
                  c->keys, c->keys_alt, c->vals, c->vals_alt, c->sort_tmp, c->left, c->right,
                  c->first, c->last, c->parent, c->lnk, c->visit, c->front, c->bbmin, c->bbmax, c->geo_c,
                  c->geo_h, c->mp, c->gstart, c->grp_first, c->grp_count, c->grp_node, c->nsel,
                  c->sel_tmp, c->stacks, c->grav_nc,
                  c->grav_nn, c->tmpd, c->grp_root, c->stacks2, c->glist, c->grp_gmac};
  for (void* p : ptrs) if (p) hipFree(p);
  c->iorig = c->iorig_alt = nullptr; c->posm = c->hrecA = c->hrecB = c->grec = nullptr;
  c->status = nullptr; c->keys = c->keys_alt = nullptr; c->vals = c->vals_alt = nullptr;
  c->sort_tmp = nullptr; c->left = c->right = c->first = c->last = c->parent = c->visit = c->front = nullptr;
  c->lnk = nullptr;
  c->bbmin = c->bbmax = nullptr; c->geo_c = c->geo_h = nullptr; c->mp = nullptr;
  c->gstart = c->grp_first = c->grp_count = c->grp_node = c->nsel = nullptr; c->sel_tmp = nullptr;
  c->stacks = nullptr;
  sph_free_lists(c);
  c->grav_nc = c->grav_nn = nullptr; c->tmpd = nullptr; c->grp_root = nullptr;
  c->grp_gmac = nullptr;
  c->stacks2 = nullptr; c->glist = nullptr;
  c->Ncap = 0; c->N = 0;
}

int alloc_particles(gdf_ctx* c, int N) {
  free_particles(c);
  c->Ncap = N;
  const size_t n = (size_t) N;
  for (int f = 0; f < F_NPERSIST; f++) {
    GDF_HIP(c, hipMalloc(&c->pf[f], sizeof(double)*n));
    GDF_HIP(c, hipMalloc(&c->pf_alt[f], sizeof(double)*n));
  }
  for (int f = 0; f < D_NDERIVED; f++) {
    GDF_HIP(c, hipMalloc(&c->df[f], sizeof(double)*n));
    if (c->nlevels > 1) GDF_HIP(c, hipMalloc(&c->df_alt[f], sizeof(double)*n));
  }
  for (int f = 0; f < I_NINT; f++) {
    GDF_HIP(c, hipMalloc(&c->pi[f], sizeof(int)*n));
    GDF_HIP(c, hipMalloc(&c->pi_alt[f], sizeof(int)*n));
  }
  GDF_HIP(c, hipMalloc(&c->iorig, sizeof(int)*n));
  GDF_HIP(c, hipMalloc(&c->iorig_alt, sizeof(int)*n));
  GDF_HIP(c, hipMalloc(&c->posm, sizeof(double4)*n));
  GDF_HIP(c, hipMalloc(&c->grec, sizeof(double4)*n));
  GDF_HIP(c, hipMalloc(&c->grav_nc, sizeof(int32_t)*n));
  GDF_HIP(c, hipMalloc(&c->grav_nn, sizeof(int32_t)*n));
  GDF_HIP(c, hipMalloc(&c->tmpd, sizeof(double)*3*n));
  GDF_TRY(tree_alloc(c));
  GDF_TRY(sph_alloc_lists(c));
  return GDF_OK;
}

int upload_ewald_table(gdf_ctx* c) {
  PhysParams& p = c->phys;
  std::vector<double> field;
  const int rc = ewald_build_table(p.boxsize, c->cfg.gr_bhewaldseriesn, c->cfg.n_ewald_grid,
                                   c->cfg.ewald_mult, c->cfg.efratio, field, p.eng, p.edI);
  if (rc != GDF_OK) return fail(c, rc, "Ewald table: bad parameters");
  const size_t n = field.size()/4;
  std::vector<float> ff(field.size());
  for (size_t i = 0; i < field.size(); i++) ff[i] = (float) field[i];
  GDF_HIP(c, hipMalloc(&c->ewald_tab, sizeof(double4)*n));
  GDF_HIP(c, hipMalloc(&c->ewald_tabf, sizeof(float4)*n));
  GDF_HIP(c, hipMemcpy(c->ewald_tab, field.data(), sizeof(double4)*n, hipMemcpyHostToDevice));
  GDF_HIP(c, hipMemcpy(c->ewald_tabf, ff.data(), sizeof(float4)*n, hipMemcpyHostToDevice));
  p.etab = c->ewald_tab;
  p.etabf = c->ewald_tabf;
  return GDF_OK;
}

int upload_kernel_tables(gdf_ctx* c) {
  // SmoothingKernel.h:583-604 tables: 6 linear in s, 3 in s^2; values from the oracle-free
  // host evaluation of the same polynomials (evaluated here in double)
  const int res = 1000;
  KernelDesc& k = c->phys.kern;
  std::vector<double> t(9*res);
  auto eval = [&](int w, double s) -> double {
    const bool m4 = k.type == 0;
    const double nrm = k.norm;
    // host copies of the device polynomials
    if (m4) {
      switch (w) {
        case 0: return s < 1 ? nrm*(1.0 - 1.5*s*s + 0.75*s*s*s) : (s < 2 ? 0.25*nrm*pow(2.0 - s, 3) : 0.0);
        case 1: return s < 1 ? nrm*(-3.0*s + 2.25*s*s) : (s < 2 ? -0.75*nrm*(2.0 - s)*(2.0 - s) : 0.0);
        case 2: return s < 1 ? nrm*(-3 + 1.5*5.0*s*s - 0.75*6.0*pow(s, 3))
                             : (s < 2 ? nrm*(-6.0 + 3.0*4.0*s - 1.5*5.0*s*s + 0.25*6.0*pow(s, 3)) : 0.0);
        case 3: return s < 1 ? 1.4 - 2.0*s*s + 1.5*pow(s, 4) - 0.6*pow(s, 5)
                             : (s < 2 ? 1.6 - 4.0*s*s + 4.0*pow(s, 3) - 1.5*pow(s, 4) + 0.2*pow(s, 5) : 0.0);
        case 4: return s < 1 ? 1.333333333333333333333*s - 1.2*pow(s, 3) + 0.5*pow(s, 4)
                             : (s < 2 ? 2.6666666666666666667*s - 3.0*s*s + 1.2*pow(s, 3) -
                                        0.166666666666666666667*pow(s, 4) - 0.06666666666666666667/(s*s)
                                      : 1.0/(s*s));
        default: return s < 1 ? 1.4 - 0.666666666666666666666666*s*s + 0.3*pow(s, 4) - 0.1*pow(s, 5)
                              : (s < 2 ? -1.0/(15.0*s) + 1.6 - 1.33333333333333333333333333*s*s + pow(s, 3) -
                                         0.3*pow(s, 4) + (1.0/30.0)*pow(s, 5)
                                       : 1.0/s);
      }
    }
    const double c12 = 12.0/359.0;
    switch (w) {
      case 0: return s < 1 ? nrm*(66.0 - 60.0*s*s + 30.0*pow(s, 4) - 10.0*pow(s, 5))
                   : s < 2 ? nrm*(51.0 + 75.0*s - 210.0*s*s + 150.0*pow(s, 3) - 45.0*pow(s, 4) + 5.0*pow(s, 5))
                   : s < 3 ? nrm*(243.0 - 405*s + 270.0*s*s - 90.0*pow(s, 3) + 15.0*pow(s, 4) - pow(s, 5)) : 0.0;
      case 1: return s < 1 ? nrm*(-120.0*s + 120.0*pow(s, 3) - 50.0*pow(s, 4))
                   : s < 2 ? nrm*(75.0 - 420.0*s + 450.0*s*s - 180.0*pow(s, 3) + 25.0*pow(s, 4))
                   : s < 3 ? nrm*(-405.0 + 540.0*s - 270.0*s*s + 60.0*pow(s, 3) - 5.0*pow(s, 4)) : 0.0;
      case 2: return s < 1 ? nrm*(-198.0 + 60.0*5.0*s*s - 30.0*7.0*pow(s, 4) + 10.0*8.0*pow(s, 5))
                   : s < 2 ? nrm*(-153.0 - 75.0*4.0*s + 210.0*5.0*s*s - 150.0*6.0*pow(s, 3) + 45.0*7.0*pow(s, 4) - 5.0*8.0*pow(s, 5))
                   : s < 3 ? nrm*(-729.0 + 405.0*4.0*s - 270.0*5.0*s*s + 90.0*6.0*pow(s, 3) - 15.0*7.0*pow(s, 4) + 8.0*pow(s, 5)) : 0.0;
      case 3: return s < 1 ? 33.0*s*s - 15.0*pow(s, 4) + 5.0*pow(s, 6) - 1.42857142857*pow(s, 7) - 34.14285714
                   : s < 2 ? 25.5*s*s + 25.0*pow(s, 3) - 52.5*pow(s, 4) + 30.0*pow(s, 5) - 7.5*pow(s, 6) + 0.7142857143*pow(s, 7) - 33.785714286
                   : s < 3 ? 121.5*s*s - 135.0*pow(s, 3) + 67.5*pow(s, 4) - 18.0*pow(s, 5) + 2.5*pow(s, 6) - 0.142857143*pow(s, 7) - 52.07142857 : 0.0;
      case 4: return s < 1 ? c12*(22.0*s - 12.0*pow(s, 3) + (30.0/7.0)*pow(s, 5) - (5.0/4.0)*pow(s, 6))
                   : s < 2 ? c12*(17.0*s + (75.0/4.0)*s*s - 42.0*pow(s, 3) + 25.0*pow(s, 4) - (45.0/7.0)*pow(s, 5) + (5.0/8.0)*pow(s, 6) + (5.0/56.0)/(s*s))
                   : s < 3 ? c12*(81.0*s - (405.0/4.0)*pow(s, 2) + 54.0*pow(s, 3) - 15.0*pow(s, 4) + (15.0/7.0)*pow(s, 5) - (1.0/8.0)*pow(s, 6) - (507.0/56.0)/(s*s))
                   : 1.0/(s*s);
      default: return s < 1 ? c12*(-11.0*s*s + 3.0*pow(s, 4) - (5.0/7.0)*pow(s, 6) + (5.0/28.0)*pow(s, 7) + (478.0/14.0))
                    : s < 2 ? c12*(-(17.0/2.0)*s*s - (25.0/4.0)*pow(s, 3) + (21.0/2.0)*pow(s, 4) - 5.0*pow(s, 5) + (15.0/14.0)*pow(s, 6) - (5.0/56.0)*pow(s, 7) + (473.0/14.0) + (5.0/56.0)/s)
                    : s < 3 ? c12*(-(81.0/2.0)*s*s + (135.0/4.0)*pow(s, 3) - (27.0/2.0)*pow(s, 4) + 3.0*pow(s, 5) - (5.0/14.0)*pow(s, 6) + (1.0/56.0)*pow(s, 7) + (729.0/14.0) - (507.0/56.0)/s)
                    : 1.0/s;
    }
  };
  const double step = k.range/res, stepsq = k.rangesqd/res;
  for (int w = 0; w < 6; w++) for (int i = 0; i < res; i++) t[w*res + i] = eval(w, step*i);
  const int sqw[3] = {0, 2, 3};
  for (int w = 0; w < 3; w++) for (int i = 0; i < res; i++) t[(6 + w)*res + i] = eval(sqw[w], sqrt(stepsq*i));
  std::vector<float> tf(t.begin(), t.end());
  GDF_HIP(c, hipMalloc(&c->ktab, sizeof(double)*9*res));
  GDF_HIP(c, hipMalloc(&c->ktabf, sizeof(float)*9*res));
  GDF_HIP(c, hipMemcpy(c->ktab, t.data(), sizeof(double)*9*res, hipMemcpyHostToDevice));
  GDF_HIP(c, hipMemcpy(c->ktabf, tf.data(), sizeof(float)*9*res, hipMemcpyHostToDevice));
  k.tab = c->ktab; k.tabf = c->ktabf; k.res = res;
  return GDF_OK;
}

}  // namespace

// context helpers for the serving pool (gdf_pool.hip)
int ctx_alloc_particles(gdf_ctx* c, int ncap) { return alloc_particles(c, ncap); }

// a field of particles [i0, i0 + n) whose iorig entries number them 0..n-1 (one segment of a
// serving batch), scattered into that local order
int ctx_get_field_range(gdf_ctx* c, const char* name, int i0, int n, double* out) {
  double* p[3] = {nullptr, nullptr, nullptr};
  int w = 0;
  if (!field_ptrs(c, name, p, &w)) return fail(c, GDF_ERR_INVALID, std::string("unknown field ") + name);
  if (n <= 0) return GDF_OK;
  k_scatter_out<<<nblocks(n, BS), BS, 0, c->stream>>>(n, w, p[0] + i0, p[1] ? p[1] + i0 : nullptr,
                                                      p[2] ? p[2] + i0 : nullptr, c->iorig + i0,
                                                      c->tmpd);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipMemcpyAsync(out, c->tmpd, sizeof(double)*w*(size_t) n, hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int ctx_field_width(gdf_ctx* c, const char* name) {
  double* p[3] = {nullptr, nullptr, nullptr};
  int w = 0;
  return field_ptrs(c, name, p, &w) ? w : 0;
}

}  // namespace gdf

using namespace gdf;

namespace gdf {

int ktime_begin(gdf_ctx* c, hipStream_t s, int kind) {
  if (!c->timing) return -1;
  if (c->nkev == gdf_ctx::kKevPairs) ktime_collect(c);
  const int k = c->nkev++;
  c->kev_kind[k] = kind;
  hipEventRecord(c->kev[2*k], s);
  return k;
}

void ktime_end(gdf_ctx* c, hipStream_t s, int slot) {
  if (slot >= 0) hipEventRecord(c->kev[2*slot + 1], s);
}

void ktime_collect(gdf_ctx* c) {
  for (int k = 0; k < c->nkev; k++) {
    float ms = 0.f;
    if (hipEventSynchronize(c->kev[2*k + 1]) == hipSuccess &&
        hipEventElapsedTime(&ms, c->kev[2*k], c->kev[2*k + 1]) == hipSuccess) {
      if (c->kev_kind[k] == 5) { c->pass_ms[5] += ms; c->grav_launches++; }
      else { c->pass_ms[6] += ms; c->hydro_launches++; }
    }
  }
  c->nkev = 0;
}

}  // namespace gdf

extern "C" {

int gdf_config_default(gdf_config* cfg) {
  if (!cfg) return GDF_ERR_INVALID;
  memset(cfg, 0, sizeof(*cfg));
  cfg->kernel = GDF_KERNEL_M4;
  cfg->tabulated_kernel = 1;
  cfg->h_fac = 1.2; cfg->h_converge = 0.01;
  cfg->gas_eos = GDF_EOS_ENERGY_EQN;
  cfg->gamma_eos = 1.66666666666666; cfg->temp0 = 1.0; cfg->mu_bar = 1.0; cfg->rho_bary = 1.0e-14;
  cfg->alpha_visc = 1.0; cfg->beta_visc = 2.0;
  cfg->hydro_forces = 1; cfg->self_gravity = 0;
  cfg->thetamaxsqd = 0.1; cfg->Nleafmax = 6; cfg->periodic = 0;
  for (int k = 0; k < 3; k++) { cfg->boxmin[k] = -9.9e30; cfg->boxmax[k] = 9.9e30; }
  cfg->courant_mult = 0.15; cfg->accel_mult = 0.3; cfg->energy_mult = 0.4;
  cfg->precision = GDF_PREC_FP64;
  cfg->group_max = 64;
  cfg->nbody = 0; cfg->nbody_direct = 0; cfg->nbody_eps = 0.0;
  cfg->nbody_mult = 0.1; cfg->fixed_dt = 0.0;
  cfg->Nlevels = 1; cfg->level_diff_max = 1;
  cfg->nbody_softening = 1; cfg->conservative_sph_star_gravity = 1;
  cfg->n_ewald_grid = 16; cfg->gr_bhewaldseriesn = 10; cfg->ewald_mult = 1.0;
  cfg->efratio = 1.0;
  cfg->ntreebuildstep = 1; cfg->ntreestockstep = 1;
  cfg->graph_capture = 0;
  cfg->multipole = GDF_MULTIPOLE_QUADRUPOLE;
  cfg->gravity_mac = GDF_MAC_GEOMETRIC; cfg->macerror = 1.0e-4;
  return GDF_OK;
}

int gdf_create(const gdf_config* cfg, gdf_ctx** out) {
  if (!cfg || !out) return GDF_ERR_INVALID;
  *out = nullptr;
  if (cfg->kernel != GDF_KERNEL_M4 && cfg->kernel != GDF_KERNEL_QUINTIC) return GDF_ERR_INVALID;
  if (cfg->gas_eos < 0 || cfg->gas_eos > 2) return GDF_ERR_INVALID;
  if (cfg->group_max < 1 || cfg->group_max > kMaxGroup) return GDF_ERR_INVALID;
  if (cfg->Nleafmax < 1 || cfg->Nleafmax > 14) return GDF_ERR_INVALID;
  if (!(cfg->thetamaxsqd > 0.0) || !(cfg->h_fac > 0.0)) return GDF_ERR_INVALID;
  if (cfg->nbody_direct < 0 || cfg->nbody_direct > 2 || cfg->nbody_eps < 0.0 ||
      cfg->fixed_dt < 0.0)
    return GDF_ERR_INVALID;
  if (cfg->Nlevels < 1 || cfg->Nlevels > 20 || cfg->level_diff_max < 1) return GDF_ERR_INVALID;
  if (cfg->hrho_iteration_max < 0 || cfg->hrho_iteration_max > 1000 ||
      cfg->cand_batches_init < 0 || cfg->cand_batches_init > 4096 || !(cfg->h_converge >= 0.0))
    return GDF_ERR_INVALID;
  if (cfg->Nlevels > 1 && (cfg->nbody || cfg->fixed_dt > 0.0)) return GDF_ERR_INVALID;
  if (cfg->periodic && cfg->self_gravity && !cfg->nbody &&
      (cfg->n_ewald_grid < 4 || cfg->n_ewald_grid > 256 || cfg->gr_bhewaldseriesn < 1 ||
       cfg->gr_bhewaldseriesn > 64 || !(cfg->ewald_mult > 0.0) || !(cfg->efratio >= 1.0) ||
       !(cfg->efratio <= 1.5)))
    return GDF_ERR_INVALID;
  if (cfg->ntreebuildstep < 1 || cfg->ntreestockstep < 1 ||
      (cfg->graph_capture != 0 && cfg->graph_capture != 1) ||
      cfg->multipole < GDF_MULTIPOLE_QUADRUPOLE || cfg->multipole > GDF_MULTIPOLE_FAST_QUADRUPOLE ||
      cfg->gravity_mac < GDF_MAC_GEOMETRIC || cfg->gravity_mac > GDF_MAC_EIGENMAC ||
      !(cfg->macerror > 0.0))
    return GDF_ERR_INVALID;
  if ((cfg->nbody_softening != 0 && cfg->nbody_softening != 1) ||
      (cfg->conservative_sph_star_gravity != 0 && cfg->conservative_sph_star_gravity != 1))
    return GDF_ERR_INVALID;
  if (cfg->periodic)
    for (int k = 0; k < 3; k++)
      if (!(cfg->boxmax[k] > cfg->boxmin[k]) || cfg->boxmax[k] - cfg->boxmin[k] > 1e20)
        return GDF_ERR_INVALID;
  gdf_ctx* c = new (std::nothrow) gdf_ctx();
  if (!c) return GDF_ERR_NOMEM;
  c->cfg = *cfg;
  if (hipGetDevice(&c->device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return GDF_ERR_HIP;
  }
  PhysParams& p = c->phys;
  KernelDesc& k = p.kern;
  k.type = cfg->kernel; k.tabulated = cfg->tabulated_kernel;
  if (k.type == 0) { k.range = 2.0; k.rangesqd = 4.0; k.norm = 0.31830988618379; }
  else { k.range = 3.0; k.rangesqd = 9.0; k.norm = 0.31830988618379*(1/120.); }
  k.res = 1000; k.resinv = 1000/k.range; k.resinvsqd = 1000/k.rangesqd;
  p.h_fac = cfg->h_fac; p.h_converge = cfg->h_converge;
  p.gamma = cfg->gamma_eos; p.gammam1 = cfg->gamma_eos - 1.0;
  p.temp0 = cfg->temp0; p.mu_bar = cfg->mu_bar; p.invrho_bary = 1.0/cfg->rho_bary;
  p.eos = cfg->gas_eos; p.energy_integration = cfg->gas_eos == GDF_EOS_ENERGY_EQN;
  p.alpha_visc = cfg->alpha_visc; p.beta_visc = cfg->beta_visc;
  p.periodic = cfg->periodic;
  for (int d = 0; d < 3; d++) {
    p.boxmin[d] = cfg->boxmin[d]; p.boxmax[d] = cfg->boxmax[d];
    p.boxsize[d] = cfg->boxmax[d] - cfg->boxmin[d];
    p.boxhalf[d] = 0.5*p.boxsize[d];
  }
  p.thetamaxsqd = cfg->thetamaxsqd; p.nleaf = cfg->Nleafmax;
  p.courant_mult = cfg->courant_mult; p.accel_mult = cfg->accel_mult;
  p.energy_mult = cfg->energy_mult;
  p.hydro_forces = cfg->hydro_forces; p.self_gravity = cfg->self_gravity;
  p.nbody = cfg->nbody != 0;
  if (p.nbody) { p.hydro_forces = 0; p.self_gravity = 1; p.energy_integration = 0; }
  p.eps2 = cfg->nbody_eps*cfg->nbody_eps;
  p.nbody_mult = cfg->nbody_mult;
  p.hrho_itmax = cfg->hrho_iteration_max > 0 ? cfg->hrho_iteration_max : 30;
  p.multipole = cfg->multipole;
  p.gravity_mac = cfg->gravity_mac; p.macerror = cfg->macerror;
  c->nlevels = cfg->Nlevels; c->level_diff_max = cfg->level_diff_max;
  int rc = upload_kernel_tables(c);
  // a periodic box with self-gravity carries the Ewald correction (Simulation.cpp:1183-1198)
  p.ewald = cfg->periodic && cfg->self_gravity && !cfg->nbody;
  if (rc == GDF_OK && p.ewald) rc = upload_ewald_table(c);
  if (rc == GDF_OK) {
    if (hipMalloc(&c->counters, sizeof(int)*32) != hipSuccess ||
        hipMalloc(&c->icount, sizeof(long long)*8) != hipSuccess ||
        hipMalloc(&c->red, sizeof(double)*16384) != hipSuccess ||
        hipHostMalloc(&c->h_counters, sizeof(int)*32) != hipSuccess ||
        hipHostMalloc(&c->h_red, sizeof(double)*16384) != hipSuccess)
      rc = GDF_ERR_NOMEM;
  }
  if (rc == GDF_OK) {
    hipMemset(c->counters, 0, sizeof(int)*32);
    if (hipStreamCreate(&c->stream2) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_dens, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_grav, hipEventDisableTiming) != hipSuccess)
      rc = GDF_ERR_HIP;
    hipMemset(c->icount, 0, sizeof(long long)*8);
    for (int e = 0; e < 12; e++) hipEventCreate(&c->ev[e]);
    for (int e = 0; e < 2*gdf_ctx::kKevPairs; e++) hipEventCreate(&c->kev[e]);
  }
  if (rc != GDF_OK) { gdf_destroy(c); return rc; }
  *out = c;
  return GDF_OK;
}

void gdf_destroy(gdf_ctx* c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  free_particles(c);
  stars_free(c);
  if (c->counters) hipFree(c->counters);
  if (c->icount) hipFree(c->icount);
  if (c->red) hipFree(c->red);
  if (c->ktab) hipFree(c->ktab);
  if (c->ktabf) hipFree(c->ktabf);
  if (c->ewald_tab) hipFree(c->ewald_tab);
  if (c->ewald_tabf) hipFree(c->ewald_tabf);
  void* segp[] = {c->d_seg_off, c->seg_root, c->seg_dt, c->seg_bnd};
  for (void* q : segp) if (q) hipFree(q);
  if (c->h_counters) hipHostFree(c->h_counters);
  if (c->h_red) hipHostFree(c->h_red);
  for (int e = 0; e < 12; e++) if (c->ev[e]) hipEventDestroy(c->ev[e]);
  for (int e = 0; e < 2*gdf_ctx::kKevPairs; e++) if (c->kev[e]) hipEventDestroy(c->kev[e]);
  if (c->stream2) hipStreamSynchronize(c->stream2);
  if (c->ev_dens) hipEventDestroy(c->ev_dens);
  if (c->ev_grav) hipEventDestroy(c->ev_grav);
  if (c->stream2) hipStreamDestroy(c->stream2);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

const char* gdf_last_error(const gdf_ctx* c) {
  if (c) return c->err.c_str();
  const char* e = noalloc_last_error();     // the context-free entry point's last failure
  return e ? e : "null context";
}

int gdf_num_particles(const gdf_ctx* c) { return c ? c->N : 0; }

int gdf_set_particles(gdf_ctx* c, int N, const double* r, const double* v, const double* m,
                      const double* h, const double* u) {
  if (!c) return GDF_ERR_INVALID;
  if (N <= 0 || !r || !v || !m || !u) return fail(c, GDF_ERR_INVALID, "gdf_set_particles: bad arguments");
  if ((long long) N > (1LL << 30)) return fail(c, GDF_ERR_INVALID, "too many particles");
  c->snap_porig.clear();
  for (int i = 0; i < N; i++)
    if (!(m[i] > 0.0)) return fail(c, GDF_ERR_INVALID, "particle masses must be positive");
  if (N > c->Ncap) GDF_TRY(alloc_particles(c, N));
  c->N = N;
  const size_t n = (size_t) N;
  std::vector<double> col(n);
  auto up = [&](double* dst, const double* src, int stride, int off) -> int {
    for (size_t i = 0; i < n; i++) col[i] = src ? src[i*stride + off] : 0.0;
    GDF_HIP(c, hipMemcpy(dst, col.data(), sizeof(double)*n, hipMemcpyHostToDevice));
    return GDF_OK;
  };
  for (int k = 0; k < 3; k++) {
    GDF_TRY(up(c->pf[F_X + k], r, 3, k)); GDF_TRY(up(c->pf[F_X0 + k], r, 3, k));
    GDF_TRY(up(c->pf[F_VX + k], v, 3, k)); GDF_TRY(up(c->pf[F_VX0 + k], v, 3, k));
    GDF_TRY(up(c->pf[F_AX0 + k], nullptr, 1, 0));
  }
  GDF_TRY(up(c->pf[F_M], m, 1, 0));
  GDF_TRY(up(c->pf[F_H], h, 1, 0));
  GDF_TRY(up(c->pf[F_U], u, 1, 0));
  GDF_TRY(up(c->pf[F_U0], u, 1, 0));
  GDF_TRY(up(c->pf[F_DUDT0], nullptr, 1, 0));
  for (int f : {F_TLAST, F_PDT, F_DTNEXT}) GDF_TRY(up(c->pf[f], nullptr, 1, 0));
  for (int f = 0; f < D_NDERIVED; f++) GDF_HIP(c, hipMemset(c->df[f], 0, sizeof(double)*n));
  for (int f = 0; f < I_NINT; f++) GDF_HIP(c, hipMemset(c->pi[f], 0, sizeof(int)*n));
  k_iota<<<nblocks(N, BS), BS, 0, c->stream>>>(N, c->iorig);
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  c->tree_valid = false; c->topo_valid = false; c->sg_valid = false; c->sg_union = false; c->setup_done = false;
  c->t = c->timestep = c->tlast = 0.0; c->nsteps = 0;
  return GDF_OK;
}

int gdf_get_field(gdf_ctx* c, const char* name, double* out) {
  if (!c || !name || !out) return GDF_ERR_INVALID;
  if (c->N <= 0) return fail(c, GDF_ERR_INVALID, "no particles");
  double* p[3] = {nullptr, nullptr, nullptr};
  int w = 0;
  const int fi = int_field(name);
  if (fi >= 0) {
    k_scatter_out_int<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, c->pi[fi], c->iorig, c->tmpd);
    GDF_HIP(c, hipGetLastError());
    GDF_HIP(c, hipMemcpyAsync(out, c->tmpd, sizeof(double)*(size_t) c->N, hipMemcpyDeviceToHost, c->stream));
    GDF_HIP(c, hipStreamSynchronize(c->stream));
    return GDF_OK;
  }
  if (!field_ptrs(c, name, p, &w)) return fail(c, GDF_ERR_INVALID, std::string("unknown field ") + name);
  k_scatter_out<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, w, p[0], p[1], p[2], c->iorig, c->tmpd);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipMemcpyAsync(out, c->tmpd, sizeof(double)*w*(size_t) c->N, hipMemcpyDeviceToHost, c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_set_field(gdf_ctx* c, const char* name, const double* in) {
  if (!c || !name || !in) return GDF_ERR_INVALID;
  if (c->N <= 0) return fail(c, GDF_ERR_INVALID, "no particles");
  double* p[3] = {nullptr, nullptr, nullptr};
  int w = 0;
  const int fi = int_field(name);
  if (fi >= 0) {
    GDF_HIP(c, hipMemcpyAsync(c->tmpd, in, sizeof(double)*(size_t) c->N, hipMemcpyHostToDevice,
                              c->stream));
    k_gather_in_int<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, c->pi[fi], c->iorig, c->tmpd);
    GDF_HIP(c, hipGetLastError());
    GDF_HIP(c, hipStreamSynchronize(c->stream));
    return GDF_OK;
  }
  if (!field_ptrs(c, name, p, &w)) return fail(c, GDF_ERR_INVALID, std::string("unknown field ") + name);
  GDF_HIP(c, hipMemcpyAsync(c->tmpd, in, sizeof(double)*w*(size_t) c->N, hipMemcpyHostToDevice, c->stream));
  k_gather_in<<<nblocks(c->N, BS), BS, 0, c->stream>>>(c->N, w, p[0], p[1], p[2], c->iorig, c->tmpd);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  if (!strcmp(name, "r") || !strcmp(name, "m") || !strcmp(name, "h")) c->tree_valid = false;
  c->sg_valid = false; c->sg_union = false;
  return GDF_OK;
}

int gdf_build_tree(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  return tree_build(c);
}

int gdf_update_all_sph_properties(gdf_ctx* c, int nopbc) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (!c->tree_valid) GDF_TRY(tree_build(c));
  return sph_density(c, nopbc);
}

int gdf_update_all_sph_hydro_forces(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "hydro forces before the density pass");
  GDF_TRY(integ_zero_accelerations(c));
  GDF_TRY(sph_build_sg(c));
  GDF_TRY(sph_hydro(c, 0));
  return stars_on_gas(c);
}

int gdf_update_all_sph_forces(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "forces before the density pass");
  GDF_TRY(integ_zero_accelerations(c));
  if (c->phys.hydro_forces) {
    GDF_TRY(gravity_launch(c));        // beside the hydro pass, on the second stream
    GDF_TRY(sph_build_sg(c));
    GDF_TRY(sph_hydro(c, 1));
    GDF_TRY(gravity_join(c));
    return stars_on_gas(c);
  }
  GDF_TRY(gravity_walk(c));
  GDF_TRY(integ_add_atree(c));
  return stars_on_gas(c);
}

int gdf_graph_state(const gdf_ctx* c, int64_t* replays, int* in_use) {
  if (!c) return GDF_ERR_INVALID;
  if (replays) *replays = c->graph_replays;
  if (in_use) *in_use = !c->cfg.graph_capture ? 0 : (c->graph_failed ? -c->graph_why : 1);
  return GDF_OK;
}

int gdf_build_tree_step(gdf_ctx* c, int rebuild_tree, int n, int ntreebuildstep,
                        int ntreestockstep) {
  if (!c || c->N <= 0 || ntreebuildstep < 1 || ntreestockstep < 1) return GDF_ERR_INVALID;
  if (rebuild_tree || !c->topo_valid || n%ntreebuildstep == 0) {
    GDF_TRY(tree_build(c));
    return 1;
  }
  GDF_TRY(tree_restock(c));
  return 0;
}

int gdf_set_opening_criterion(gdf_ctx* c, int mac) {
  if (!c || mac < GDF_MAC_GEOMETRIC || mac > GDF_MAC_EIGENMAC) return GDF_ERR_INVALID;
  c->phys.gravity_mac = mac;
  c->cfg.gravity_mac = mac;
  return GDF_OK;
}

int gdf_opening_criterion(const gdf_ctx* c) {
  if (!c) return GDF_ERR_INVALID;
  return c->phys.gravity_mac;
}

int gdf_maximum_smoothing_length(gdf_ctx* c, double* hmax) {
  if (!c || !hmax || c->N <= 0) return GDF_ERR_INVALID;
  return integ_max_h(c, hmax);
}

int gdf_advance_particles(gdf_ctx* c, double dt) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  return integ_advance(c, dt);
}

int gdf_check_boundaries(gdf_ctx* c) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  return integ_wrap(c);
}

int gdf_end_timestep(gdf_ctx* c, double dtstep) {
  if (!c || c->N <= 0) return GDF_ERR_INVALID;
  GDF_TRY(integ_end_step(c, dtstep));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int gdf_compute_global_timestep(gdf_ctx* c, double* dt_min) {
  if (!c || !dt_min || c->N <= 0) return GDF_ERR_INVALID;
  return integ_timestep(c, dt_min);
}

int gdf_calculate_diagnostics(gdf_ctx* c, double* out) {
  if (!c || !out || c->N <= 0) return GDF_ERR_INVALID;
  return integ_diagnostics(c, out);
}

int gdf_time(const gdf_ctx* c, double* t, double* timestep, int64_t* nsteps) {
  if (!c) return GDF_ERR_INVALID;
  if (t) *t = c->t;
  if (timestep) *timestep = c->timestep;
  if (nsteps) *nsteps = c->nsteps;
  return GDF_OK;
}

int gdf_block_state(const gdf_ctx* c, int64_t* iout5, double* dout1) {
  if (!c) return GDF_ERR_INVALID;
  if (iout5) {
    iout5[0] = c->n_int; iout5[1] = c->level_max; iout5[2] = c->level_step;
    iout5[3] = c->nresync; iout5[4] = c->finished_last;
  }
  if (dout1) dout1[0] = c->dt_max;
  return GDF_OK;
}

int gdf_interaction_counts(gdf_ctx* c, int64_t* out6) {
  if (!c || !out6) return GDF_ERR_INVALID;
  for (int k = 0; k < 6; k++) out6[k] = c->icounts_last[k];
  return GDF_OK;
}

int gdf_gravity_group_counts(gdf_ctx* c, int32_t* ncell, int32_t* nnear) {
  if (!c || c->ngroups <= 0) return GDF_ERR_INVALID;
  if (ncell) GDF_HIP(c, hipMemcpy(ncell, c->grav_nc, sizeof(int32_t)*c->ngroups, hipMemcpyDeviceToHost));
  if (nnear) GDF_HIP(c, hipMemcpy(nnear, c->grav_nn, sizeof(int32_t)*c->ngroups, hipMemcpyDeviceToHost));
  return GDF_OK;
}

int gdf_pass_times(gdf_ctx* c, double* out9) {
  if (!c || !out9) return GDF_ERR_INVALID;
  gdf::ktime_collect(c);
  for (int k = 0; k < 6; k++) out9[k] = c->pass_ms[k];
  out9[6] = (double) c->grav_launches;
  out9[7] = c->pass_ms[6];
  out9[8] = (double) c->hydro_launches;
  return GDF_OK;
}

int gdf_set_timing(gdf_ctx* c, int enable) {
  if (!c) return GDF_ERR_INVALID;
  gdf::ktime_collect(c);
  c->timing = enable != 0;
  return GDF_OK;
}

int gdf_tree_arrays(gdf_ctx* c, int32_t* perm, uint64_t* keys, int32_t* left, int32_t* right,
                    int32_t* first, int32_t* last, double* mass, double* com, double* quad,
                    double* bbmin, double* bbmax, double* hmax, int32_t* group_first,
                    int32_t* group_count, int32_t* group_node, int* ngroups) {
  if (!c || c->N <= 0 || !c->tree_valid) return GDF_ERR_INVALID;
  const size_t N = (size_t) c->N, nn = 2*N - 1, ni = N > 1 ? N - 1 : 0;
  hipStream_t s = c->stream;
  GDF_HIP(c, hipStreamSynchronize(s));
  if (perm) GDF_HIP(c, hipMemcpy(perm, c->iorig, sizeof(int)*N, hipMemcpyDeviceToHost));
  if (keys) GDF_HIP(c, hipMemcpy(keys, c->keys, sizeof(uint64_t)*N, hipMemcpyDeviceToHost));
  if (ni) {
    if (left) GDF_HIP(c, hipMemcpy(left, c->left, sizeof(int)*ni, hipMemcpyDeviceToHost));
    if (right) GDF_HIP(c, hipMemcpy(right, c->right, sizeof(int)*ni, hipMemcpyDeviceToHost));
    if (first) GDF_HIP(c, hipMemcpy(first, c->first, sizeof(int)*ni, hipMemcpyDeviceToHost));
    if (last) GDF_HIP(c, hipMemcpy(last, c->last, sizeof(int)*ni, hipMemcpyDeviceToHost));
  }
  std::vector<double> mpv(10*nn);
  if (mass || com || quad) {
    GDF_HIP(c, hipMemcpy(mpv.data(), c->mp, sizeof(double)*10*nn, hipMemcpyDeviceToHost));
    for (size_t n = 0; n < nn; n++) {
      if (mass) mass[n] = mpv[10*n + 3];
      if (com) for (int k = 0; k < 3; k++) com[3*n + k] = mpv[10*n + k];
      if (quad) for (int k = 0; k < 5; k++) quad[5*n + k] = mpv[10*n + 4 + k];
    }
  }
  if (bbmin) GDF_HIP(c, hipMemcpy(bbmin, c->bbmin, sizeof(double)*3*nn, hipMemcpyDeviceToHost));
  if (bbmax) GDF_HIP(c, hipMemcpy(bbmax, c->bbmax, sizeof(double)*3*nn, hipMemcpyDeviceToHost));
  if (hmax) {
    std::vector<double4> gc(nn);
    GDF_HIP(c, hipMemcpy(gc.data(), c->geo_c, sizeof(double4)*nn, hipMemcpyDeviceToHost));
    for (size_t n = 0; n < nn; n++) hmax[n] = gc[n].w;
  }
  const size_t ng = (size_t) c->ngroups;
  if (group_first) GDF_HIP(c, hipMemcpy(group_first, c->grp_first, sizeof(int)*ng, hipMemcpyDeviceToHost));
  if (group_count) GDF_HIP(c, hipMemcpy(group_count, c->grp_count, sizeof(int)*ng, hipMemcpyDeviceToHost));
  if (group_node) GDF_HIP(c, hipMemcpy(group_node, c->grp_node, sizeof(int)*ng, hipMemcpyDeviceToHost));
  if (ngroups) *ngroups = c->ngroups;
  return GDF_OK;
}

}  // extern "C"

namespace {

// Per-particle sets (import order) into CSR form.
int64_t emit_sets(gdf_ctx* c, std::vector<std::vector<int>>& sets, int64_t* offsets,
                  int32_t* ids, int64_t cap) {
  const int N = c->N;
  for (auto& v : sets) std::sort(v.begin(), v.end());
  int64_t total = 0;
  offsets[0] = 0;
  for (int i = 0; i < N; i++) { total += (int64_t) sets[i].size(); offsets[i + 1] = total; }
  if (!ids) return total;
  if (total > cap) return fail(c, GDF_ERR_CAPACITY, "neighbour set output capacity too small");
  for (int i = 0; i < N; i++) std::copy(sets[i].begin(), sets[i].end(), ids + offsets[i]);
  return total;
}

// The scatter-gather sets the mixed path's last force pass evaluated: its candidate lists (the
// density pass's union lists, or the force pass's own walk) decoded with the FP32 group-frame
// test of its hydro kernel (positions relative to the group centre rounded to float, squared
// ranges rounded to float, the squared distance with two fused multiply-adds), for the pass's
// targets (its active flags in a block step, every particle otherwise): the walks bound the
// lists by the targets' box, so they need not hold the sets of a group's other members.
int64_t pass_sg_sets(gdf_ctx* c, int64_t* offsets, int32_t* ids, int64_t cap) {
  if (c->cfg.precision != GDF_PREC_MIXED || !c->sg_valid || c->sg_masks)
    return fail(c, GDF_ERR_INVALID, "no candidate lists of a mixed force pass to decode");
  const CandSet& cs = c->sg_from_union ? c->cs_g : c->cs_s;
  const int N = c->N, G = c->ngroups, nn = 2*N - 1;
  const size_t per = (size_t) cs.capb*64;
  std::vector<int> cidx((size_t) G*per), ncand(G), io(N), gfirst(G), gcount(G), gnode(G);
  std::vector<double> hh(N), x(N), y(N), z(N);
  std::vector<double4> gc(nn);
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  GDF_HIP(c, hipMemcpy(cidx.data(), cs.idx, sizeof(int)*cidx.size(), hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(ncand.data(), cs.ncand, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(io.data(), c->iorig, sizeof(int)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gfirst.data(), c->grp_first, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gcount.data(), c->grp_count, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gnode.data(), c->grp_node, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gc.data(), c->geo_c, sizeof(double4)*nn, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(hh.data(), c->pf[F_H], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(x.data(), c->pf[F_X], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(y.data(), c->pf[F_Y], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(z.data(), c->pf[F_Z], sizeof(double)*N, hipMemcpyDeviceToHost));
  std::vector<int> tgt(N, 1);
  if (!c->sg_tgt_all) {
    if (!c->sg_tgt || c->sg_tgt_cap < N)
      return fail(c, GDF_ERR_INVALID, "no target flags of the last force pass");
    GDF_HIP(c, hipMemcpy(tgt.data(), c->sg_tgt, sizeof(int)*N, hipMemcpyDeviceToHost));
  }
  const double krsqd = c->phys.kern.rangesqd;
  const double* hb = c->phys.boxhalf;
  const double* bs = c->phys.boxsize;
  auto img = [&](double& d, int k) {
    if (!c->phys.periodic) return;
    if (d > hb[k]) d -= bs[k]; else if (d < -hb[k]) d += bs[k];
  };
  auto r2f = [&](double h) { return (float) (krsqd*h*h); };
  std::vector<std::vector<int>> sets(N);
  for (int g = 0; g < G; g++) {
    const int nc = ncand[g];
    if (nc <= 0 || (size_t) nc > per) continue;
    const double4 o = gc[gnode[g]];
    const int* lst = cidx.data() + (size_t) g*per;
    std::vector<float> cx(nc), cy(nc), cz(nc), cr(nc);
    for (int k = 0; k < nc; k++) {
      const int j = lst[k];
      double dx = x[j] - o.x, dy = y[j] - o.y, dz = z[j] - o.z;
      img(dx, 0); img(dy, 1); img(dz, 2);
      cx[k] = (float) dx; cy[k] = (float) dy; cz[k] = (float) dz; cr[k] = r2f(hh[j]);
    }
    for (int l = 0; l < gcount[g]; l++) {
      const int i = gfirst[g] + l;
      if (!tgt[i]) continue;
      const float rx = (float) (x[i] - o.x), ry = (float) (y[i] - o.y), rz = (float) (z[i] - o.z);
      const float r2i = r2f(hh[i]);
      auto& out = sets[io[i]];
      for (int k = 0; k < nc; k++) {
        const float a = cx[k] - rx, b = cy[k] - ry, d = cz[k] - rz;
        const float d2 = std::fma(d, d, std::fma(b, b, a*a));
        if (d2 < r2i || d2 < cr[k]) out.push_back(io[lst[k]]);
      }
    }
  }
  return emit_sets(c, sets, offsets, ids, cap);
}

}  // namespace

extern "C" {

int64_t gdf_neighbour_sets(gdf_ctx* c, int mode, int64_t* offsets, int32_t* ids, int64_t cap) {
  if (!c || !offsets || c->N <= 0 || mode < 0 || mode > 2) return GDF_ERR_INVALID;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "no tree");
  if (mode == 2) return pass_sg_sets(c, offsets, ids, cap);
  std::vector<int> cidx, ncand;
  std::vector<unsigned long long> cmask;
  int capb = 0;
  GDF_TRY(sph_copy_sg(c, cidx, cmask, ncand, capb));
  const int N = c->N, G = c->ngroups;
  std::vector<int> io(N), gfirst(G), gcount(G);
  std::vector<double> hh(N), x(N), y(N), z(N);
  GDF_HIP(c, hipMemcpy(io.data(), c->iorig, sizeof(int)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gfirst.data(), c->grp_first, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(gcount.data(), c->grp_count, sizeof(int)*G, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(hh.data(), c->pf[F_H], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(x.data(), c->pf[F_X], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(y.data(), c->pf[F_Y], sizeof(double)*N, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(z.data(), c->pf[F_Z], sizeof(double)*N, hipMemcpyDeviceToHost));
  // the scatter-gather candidate masks hold exactly the mode-1 set; the gather set is its
  // subset with r2 < k^2 h_i^2, decided with the same arithmetic as the device tests
  const double krsqd = c->phys.kern.rangesqd;
  std::vector<std::vector<int>> sets(N);
  for (int g = 0; g < G; g++) {
    const size_t base = (size_t) g*capb*64;
    const int nc = ncand[g];
    const int nbat = (nc + 63)/64;
    for (int l = 0; l < gcount[g]; l++) {
      const int s = gfirst[g] + l;
      const double r2i = (krsqd*hh[s])*hh[s];
      auto& out = sets[io[s]];
      // candidate k (walk order) sits in batch k % nbat at bit k / nbat
      for (int k = 0; k < nc; k++) {
        const int b = k%nbat, q = k/nbat;
        if (!((cmask[base + (size_t) b*64 + l] >> q) & 1ULL)) continue;
        const int j = cidx[base + (size_t) k];
        if (mode == 0) {
          double dx = x[j] - x[s], dy = y[j] - y[s], dz = z[j] - z[s];
          if (c->phys.periodic) {
            const double* hb = c->phys.boxhalf; const double* bs = c->phys.boxsize;
            if (dx > hb[0]) dx -= bs[0]; else if (dx < -hb[0]) dx += bs[0];
            if (dy > hb[1]) dy -= bs[1]; else if (dy < -hb[1]) dy += bs[1];
            if (dz > hb[2]) dz -= bs[2]; else if (dz < -hb[2]) dz += bs[2];
          }
          const double d2 = (dx*dx + dy*dy) + dz*dz;
          if (!(d2 < r2i)) continue;
        }
        out.push_back(io[j]);
      }
      std::sort(out.begin(), out.end());
    }
  }
  int64_t total = 0;
  offsets[0] = 0;
  for (int i = 0; i < N; i++) { total += (int64_t) sets[i].size(); offsets[i + 1] = total; }
  if (!ids) return total;
  if (total > cap) return fail(c, GDF_ERR_CAPACITY, "neighbour set output capacity too small");
  for (int i = 0; i < N; i++)
    std::copy(sets[i].begin(), sets[i].end(), ids + offsets[i]);
  return total;
}

}  // extern "C"
