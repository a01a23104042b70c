// Serving pool: independent simulations of one configuration class batched in one device context.
//
// No reference counterpart: GANDALF runs one simulation per process.  A request is set up alone
// (SphSimulation::PostInitialConditionsSetup, src/Hydrodynamics/SphSimulation.cpp:204-565, or
// NbodySimulation::PostInitialConditionsSetup, src/Nbody/NbodySimulation.cpp:183-250) in a
// scratch context whose tree already uses the batch's segment-prefixed keys, then its particle
// state joins the batch as one more contiguous segment.  Every batch step advances all resident
// requests by one of their own global timesteps (SphSimulation::MainLoop, :574-880, per segment)
// with one launch of each pass over the combined walk-group list: the Morton sort keeps segments
// apart, each segment is one subtree of the LBVH, and every walk starts at its group's segment
// root, so no interaction crosses requests and each request's arithmetic does not depend on the
// others.  Requests that complete their steps leave the batch at the end of the step (their
// final fields are kept on the host) and the remaining segments are compacted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "gdf_internal.h"

namespace gdf {
namespace {

struct PoolReq {
  int id, N, nsteps, done;
  double t, timestep, tlast;
};

struct PoolResult {
  int N, steps;
  double t;
  std::map<std::string, std::vector<double>> f;
};

// fields kept for a finished request
const char* kKeep[] = {"r", "v", "a", "m", "h", "rho", "u", "dudt", "div_v", "gpot", "atree"};

}  // namespace
}  // namespace gdf

struct gdf_pool {
  gdf_config cfg;
  gdf_ctx* batch = nullptr;
  gdf_ctx* scratch = nullptr;               // where requests are set up alone (kept between
                                            // admissions: its arrays only grow)
  int64_t max_particles = 0;
  std::vector<gdf::PoolReq> res;            // resident, in segment order
  std::map<int, gdf::PoolResult> fin;       // finished
  int next_id = 0;
  std::string err;
  // set when a failure left the resident requests' state unusable (an allocation that could
  // not be undone, or a step that stopped half way): every later admit and step is refused
  bool failed = false;
};

namespace gdf {
namespace {

int pool_fail(gdf_pool* P, int code, const std::string& msg) {
  P->err = msg;
  return code;
}

int pool_ctx_fail(gdf_pool* P, gdf_ctx* c, int rc) {
  P->err = gdf_last_error(c);
  return rc;
}

// Fault injection for the failure-path tests: with GDF_FAULT_POOL_GROW=1 in the environment,
// growing a batch reports an allocation failure after the grown arrays were allocated (the
// state a failed hipMalloc part way through leaves behind).
bool grow_fault_injected() {
  const char* e = std::getenv("GDF_FAULT_POOL_GROW");
  return e && e[0] == '1';
}

// Grow the batch's particle arrays to hold `need` particles, keeping the resident state.  If
// the grown arrays cannot be allocated, the arrays are made again at the previous capacity and
// the resident state is copied back: the call fails with the batch as it was (*lost = false).
// Only if that fails too is the resident state gone (*lost = true).
int batch_reserve(gdf_ctx* c, int need, int64_t maxp, bool* lost) {
  *lost = false;
  if (need <= c->Ncap) return GDF_OK;
  int ncap = std::max(need, c->Ncap + c->Ncap/2);
  if ((int64_t) ncap > maxp) ncap = (int) std::max<int64_t>(need, maxp);
  const int N = c->N, old_cap = c->Ncap;
  double* keep[F_NPERSIST];
  for (int f = 0; f < F_NPERSIST; f++) { keep[f] = c->pf[f]; c->pf[f] = nullptr; }
  int* keep_io = c->iorig;
  c->iorig = nullptr;
  auto release_kept = [&]() {
    for (int f = 0; f < F_NPERSIST; f++) if (keep[f]) (void) hipFree(keep[f]);
    if (keep_io) (void) hipFree(keep_io);
  };
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    release_kept();
    *lost = true;
    return fail(c, GDF_ERR_HIP, "batch grow: stream synchronisation failed");
  }
  int rc = ctx_alloc_particles(c, ncap);
  if (rc == GDF_OK && grow_fault_injected())
    rc = fail(c, GDF_ERR_NOMEM, "batch grow: injected allocation failure");
  std::string why;
  if (rc != GDF_OK) {
    why = c->err;
    if (old_cap < 1 || ctx_alloc_particles(c, old_cap) != GDF_OK) {
      release_kept();
      *lost = true;
      return fail(c, rc, why + "; the resident requests could not be restored");
    }
    ncap = old_cap;
  }
  hipError_t e = hipSuccess;
  // derived fields a class never computes (density of N-body) read 0
  for (int f = 0; f < D_NDERIVED && e == hipSuccess; f++)
    e = hipMemsetAsync(c->df[f], 0, sizeof(double)*ncap, c->stream);
  if (N > 0) {
    for (int f = 0; f < F_NPERSIST && e == hipSuccess; f++)
      e = hipMemcpyAsync(c->pf[f], keep[f], sizeof(double)*N, hipMemcpyDeviceToDevice,
                         c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(c->iorig, keep_io, sizeof(int)*N, hipMemcpyDeviceToDevice, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  release_kept();
  if (e != hipSuccess) {
    *lost = true;
    return check_hip(c, e, "batch grow: copying the resident state");
  }
  c->N = N;
  // the new integer arrays need the resident segments' ids again
  if (N > 0) {
    const int rs = seg_upload(c);
    if (rs != GDF_OK) { *lost = true; return rs; }
  }
  if (rc != GDF_OK) return fail(c, rc, why + " (the resident requests are unchanged)");
  return GDF_OK;
}

// keep the listed segments (in order), closing the gaps left by the others
int batch_compact(gdf_ctx* c, const std::vector<int>& keep) {
  std::vector<int> off(1, 0);
  for (int sg : keep) off.push_back(off.back() + c->seg_off[sg + 1] - c->seg_off[sg]);
  for (int f = 0; f <= F_NPERSIST; f++) {
    double* src = f < F_NPERSIST ? c->pf[f] : nullptr;
    for (size_t k = 0; k < keep.size(); k++) {
      const int sg = keep[k], n = c->seg_off[sg + 1] - c->seg_off[sg];
      if (f < F_NPERSIST)
        GDF_HIP(c, hipMemcpyAsync(c->pf_alt[f] + off[k], src + c->seg_off[sg],
                                  sizeof(double)*n, hipMemcpyDeviceToDevice, c->stream));
      else
        GDF_HIP(c, hipMemcpyAsync(c->iorig_alt + off[k], c->iorig + c->seg_off[sg],
                                  sizeof(int)*n, hipMemcpyDeviceToDevice, c->stream));
    }
    if (f < F_NPERSIST) std::swap(c->pf[f], c->pf_alt[f]);
    else std::swap(c->iorig, c->iorig_alt);
  }
  c->N = off.back();
  c->seg_off = off;
  c->nseg = (int) keep.size();
  c->tree_valid = false;
  c->topo_valid = false;
  c->sg_valid = false; c->sg_union = false;
  if (c->nseg > 0) GDF_TRY(seg_upload(c));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

// forces of every resident request: SphSimulation::MainLoop's force pass
// (GradhSphTree::UpdateAllSphForces / UpdateAllSphHydroForces) or the N-body tree walk
int batch_forces(gdf_ctx* c) {
  if (c->phys.nbody) {
    GDF_TRY(integ_zero_forces(c));
    GDF_TRY(tree_build(c));
    GDF_TRY(gravity_walk(c));
    return integ_add_atree(c);
  }
  GDF_TRY(tree_build(c));
  GDF_TRY(sph_density(c, 0));
  GDF_TRY(integ_zero_forces(c));
  if (c->phys.self_gravity) {
    if (c->phys.hydro_forces) {
      GDF_TRY(gravity_launch(c));          // beside the hydro pass, on the second stream
      GDF_TRY(sph_build_sg(c));
      GDF_TRY(sph_hydro(c, 1));
      return gravity_join(c);
    }
    GDF_TRY(gravity_walk(c));
    return integ_add_atree(c);
  }
  GDF_TRY(sph_build_sg(c));
  return sph_hydro(c, 0);
}

int upload_seg_dt(gdf_ctx* c, const std::vector<double>& v) {
  GDF_HIP(c, hipMemcpyAsync(c->seg_dt, v.data(), sizeof(double)*v.size(), hipMemcpyHostToDevice,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

}  // namespace
}  // namespace gdf

using namespace gdf;

extern "C" {

int gdf_pool_create(const gdf_config* cfg, int64_t max_particles, gdf_pool** out) {
  if (!cfg || !out || max_particles <= 0) return GDF_ERR_INVALID;
  *out = nullptr;
  if (cfg->Nlevels != 1 || cfg->nbody_direct != 0) return GDF_ERR_INVALID;
  // serving runs several batches from several host threads: the captured tree-build graphs
  // are for a context driven by one thread alone
  if (cfg->graph_capture) return GDF_ERR_INVALID;
  if (max_particles > (int64_t) 1 << 30) return GDF_ERR_INVALID;
  gdf_pool* P = new (std::nothrow) gdf_pool();
  if (!P) return GDF_ERR_NOMEM;
  P->cfg = *cfg;
  P->max_particles = max_particles;
  int rc = gdf_create(cfg, &P->batch);
  if (rc != GDF_OK) { delete P; return rc; }
  rc = seg_alloc(P->batch);
  if (rc != GDF_OK) { gdf_destroy(P->batch); delete P; return rc; }
  *out = P;
  return GDF_OK;
}

void gdf_pool_destroy(gdf_pool* P) {
  if (!P) return;
  gdf_destroy(P->scratch);
  gdf_destroy(P->batch);
  delete P;
}

const char* gdf_pool_last_error(const gdf_pool* P) {
  if (!P) return "null pool";
  return P->failed && P->err.empty() ? "pool failed" : P->err.c_str();
}

int64_t gdf_pool_resident_particles(const gdf_pool* P) { return P ? P->batch->N : 0; }

int gdf_pool_admit(gdf_pool* P, int N, const double* r, const double* v, const double* m,
                   const double* u, int nsteps, int* id) {
  if (!P || !id || N < 2 || nsteps < 1 || !r || !v || !m || !u) return GDF_ERR_INVALID;
  if (P->failed) return GDF_ERR_INVALID;
  gdf_ctx* c = P->batch;
  if ((int64_t) c->N + N > P->max_particles || (int) P->res.size() >= kMaxSeg)
    return pool_fail(P, GDF_ERR_CAPACITY, "pool full: the request waits for resident ones to finish");
  // set the request up alone, with the batch's key layout (one segment)
  int rc = GDF_OK;
  if (!P->scratch) {
    rc = gdf_create(&P->cfg, &P->scratch);
    if (rc != GDF_OK) { P->scratch = nullptr; return pool_fail(P, rc, "scratch context creation failed"); }
  }
  gdf_ctx* sc = P->scratch;
  rc = gdf_set_particles(sc, N, r, v, m, nullptr, u);
  if (rc == GDF_OK) {
    sc->nseg = 1;
    sc->seg_off = {0, N};
    rc = seg_upload(sc);
  }
  if (rc == GDF_OK) rc = gdf_setup_simulation(sc);
  if (rc != GDF_OK) return pool_ctx_fail(P, sc, rc);
  // join the batch as the last segment
  const int N0 = c->N;
  bool lost = false;
  rc = batch_reserve(c, N0 + N, P->max_particles, &lost);
  if (lost) {
    P->failed = true;
    return pool_ctx_fail(P, c, rc);
  }
  if (rc == GDF_OK) {
    for (int f = 0; f < F_NPERSIST && rc == GDF_OK; f++)
      if (hipMemcpyAsync(c->pf[f] + N0, sc->pf[f], sizeof(double)*N, hipMemcpyDeviceToDevice,
                         c->stream) != hipSuccess) rc = GDF_ERR_HIP;
    if (rc == GDF_OK && hipMemcpyAsync(c->iorig + N0, sc->iorig, sizeof(int)*N,
                                       hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
      rc = GDF_ERR_HIP;
  }
  PoolReq q;
  q.id = P->next_id; q.N = N; q.nsteps = nsteps; q.done = 0;
  q.t = sc->t; q.timestep = sc->timestep; q.tlast = sc->tlast;
  if (rc == GDF_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = GDF_ERR_HIP;
  if (rc != GDF_OK) return pool_ctx_fail(P, c, rc);
  c->N = N0 + N;
  if (c->seg_off.empty()) c->seg_off.push_back(0);
  c->seg_off.push_back(c->N);
  c->nseg = (int) c->seg_off.size() - 1;
  rc = seg_upload(c);
  if (rc != GDF_OK) return pool_ctx_fail(P, c, rc);
  P->res.push_back(q);
  *id = P->next_id++;
  return GDF_OK;
}

int gdf_pool_step(gdf_pool* P, int64_t* out2) {
  if (!P) return GDF_ERR_INVALID;
  if (P->failed) return GDF_ERR_INVALID;
  gdf_ctx* c = P->batch;
  if (out2) { out2[0] = 0; out2[1] = 0; }
  const int ns = (int) P->res.size();
  if (ns == 0) return GDF_OK;
  // drift length t - tlast of each request (SphLeapfrogKDK::AdvanceParticles)
  std::vector<double> adv(ns), kick(ns), newdt(ns);
  for (int sg = 0; sg < ns; sg++) {
    PoolReq& q = P->res[sg];
    kick[sg] = q.timestep;
    q.t = q.t + q.timestep;
    adv[sg] = q.t - q.tlast;
  }
  int rc = upload_seg_dt(c, adv);
  if (rc == GDF_OK) rc = integ_seg_advance(c);
  if (rc == GDF_OK && !c->phys.nbody) rc = integ_wrap(c);
  if (rc == GDF_OK) rc = batch_forces(c);
  if (rc == GDF_OK) {
    if (P->cfg.fixed_dt > 0.0) std::fill(newdt.begin(), newdt.end(), P->cfg.fixed_dt);
    else rc = integ_seg_timesteps(c, newdt.data());
  }
  if (rc == GDF_OK) rc = upload_seg_dt(c, kick);
  if (rc == GDF_OK) rc = integ_seg_end_step(c);
  if (rc == GDF_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = GDF_ERR_HIP;
  if (rc != GDF_OK) {
    // the requests stopped part way through their step: their state is not a step boundary
    P->failed = true;
    return pool_ctx_fail(P, c, rc);
  }
  int64_t psteps = 0;
  std::vector<int> keep;
  for (int sg = 0; sg < ns; sg++) {
    PoolReq& q = P->res[sg];
    q.tlast = q.t;
    q.timestep = newdt[sg];
    q.done++;
    psteps += q.N;
    if (q.done < q.nsteps) { keep.push_back(sg); continue; }
    // finished: keep its final state on the host
    PoolResult pr;
    pr.N = q.N; pr.steps = q.done; pr.t = q.t;
    for (const char* name : kKeep) {
      const int w = ctx_field_width(c, name);
      std::vector<double> buf((size_t) w*q.N);
      rc = ctx_get_field_range(c, name, c->seg_off[sg], q.N, buf.data());
      if (rc != GDF_OK) return pool_ctx_fail(P, c, rc);
      pr.f[name] = std::move(buf);
    }
    P->fin[q.id] = std::move(pr);
  }
  if ((int) keep.size() < ns) {
    std::vector<PoolReq> nres;
    for (int sg : keep) nres.push_back(P->res[sg]);
    rc = batch_compact(c, keep);
    if (rc != GDF_OK) return pool_ctx_fail(P, c, rc);
    P->res.swap(nres);
  }
  if (out2) { out2[0] = ns; out2[1] = psteps; }
  return GDF_OK;
}

int gdf_pool_status(const gdf_pool* P, int id, int64_t* out3, double* t) {
  if (!P || !out3) return GDF_ERR_INVALID;
  out3[0] = 0; out3[1] = 0; out3[2] = 0;
  for (const PoolReq& q : P->res)
    if (q.id == id) {
      out3[0] = 1; out3[1] = q.done; out3[2] = q.N;
      if (t) *t = q.t;
      return GDF_OK;
    }
  auto it = P->fin.find(id);
  if (it != P->fin.end()) {
    out3[0] = 2; out3[1] = it->second.steps; out3[2] = it->second.N;
    if (t) *t = it->second.t;
  }
  return GDF_OK;
}

int gdf_pool_get_field(gdf_pool* P, int id, const char* name, double* out) {
  if (!P || !name || !out) return GDF_ERR_INVALID;
  auto it = P->fin.find(id);
  if (it != P->fin.end()) {
    auto f = it->second.f.find(name);
    if (f == it->second.f.end())
      return pool_fail(P, GDF_ERR_INVALID, std::string("field not kept for finished requests: ") + name);
    std::copy(f->second.begin(), f->second.end(), out);
    return GDF_OK;
  }
  gdf_ctx* c = P->batch;
  for (size_t sg = 0; sg < P->res.size(); sg++)
    if (P->res[sg].id == id) {
      const int rc = ctx_get_field_range(c, name, c->seg_off[sg], P->res[sg].N, out);
      return rc == GDF_OK ? rc : pool_ctx_fail(P, c, rc);
    }
  return pool_fail(P, GDF_ERR_INVALID, "unknown request id");
}

int gdf_pool_release(gdf_pool* P, int id) {
  if (!P) return GDF_ERR_INVALID;
  P->fin.erase(id);
  return GDF_OK;
}

}  // extern "C"
