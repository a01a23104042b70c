// grad-h SPH passes: candidate walks, h-rho iteration and hydro forces.
//
// Replaces the OpenMP cell loops of GradhSphTree::UpdateAllSphProperties
// (src/GradhSph/GradhSphTree.cpp:83-271) and UpdateAllSphHydroForces (:280-435) with the
// per-particle physics of GradhSph::ComputeH (src/GradhSph/GradhSph.cpp:142-326) and
// ComputeSphHydroForces / ComputeSphHydroGravForces (:361-460, :474-585).
//
// Work mapping: one wavefront per walk group (a tree node of <= 64 Morton-consecutive
// particles, one target per lane).  A walk pass resolves the group's candidate particles from
// the tree once (all particles of buckets that come within the group's search range, in walk
// order) and, for every batch of up to 64 candidates, records one 64-bit mask per target
// holding the candidates that are inside that target's own range (the exact neighbour test).
// Batch b of nbat takes the candidates at walk positions b, b + nbat, b + 2 nbat, ...: every
// batch then samples the whole candidate region, each target's neighbours spread evenly over
// the batches, and the per-target mask loops below leave fewer lanes idle than with batches of
// consecutive (spatially clustered) candidates.  The pair passes stage each batch's candidate
// records in LDS with one load per lane, shared by all 64 targets, and every lane walks the set
// bits of its own mask.  The h-rho pass re-sweeps the same batches until every target of the
// group has converged.  A target sums its neighbours batch by batch, in bit order within a
// batch.
#include <hip/hip_runtime.h>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;
constexpr double NODE_SLACK = 1.0 + 1e-12; // node tests err on the inclusive side
// the per-member filter of the candidate walks errs further on the inclusive side than any
// target's own test (FP32 tests of the mixed path included)
constexpr double KEEP_SLACK = 1.0 + 1e-4;
constexpr int kInitBatches = 16;           // initial candidate capacity per group, in batches

struct CandWalk {
  int N, ngroups, nleaf, capb;
  const int* grp_node; const int* grp_first; const int* grp_count;
  const int4* lnk;
  const double4* geo_c; const double4* geo_h;
  const double* x; const double* y; const double* z; const double* hcand; const double* htarget;
  const int* status;                 // gather walks: only targets with status 1 take part
  const int* grp_root;               // serving batches: the group's segment subtree (NULL = root)
  const int* glist;                  // groups to walk (count at glist[Ncap]); NULL = all groups
  int ncap;
  int* idx; unsigned long long* mask; int* ncand;
  int* stacks; int* counter; int* flags;
  PhysParams p;
};

// The mixed gather pass's candidate radius^2 in FP32: k hs_i widened by 1e-6 of the group's
// half-extent plus k hs_i, squared and widened by 1e-5 (r2i = (k hs_i)^2)
__device__ __forceinline__ float gather_r2f(double r2i, const double4& gh) {
#pragma clang fp contract(off)
  const double rg = sqrt(r2i) + 1.0e-6*(gh.x + gh.y + gh.z + sqrt(r2i));
  return (float) (rg*rg*(1.0 + 1.0e-5));
}

// MODE 0: search set |r_ij|^2 <= (k hs_i)^2 (FP64); 1: scatter-gather set
// |r_ij|^2 < (k h_i)^2 or < (k h_j)^2 (FP64, the oracle's operation order); 2: the same set
// decided in FP32 in the group-centred frame, exactly as the mixed-precision gravity walk
// decides it; 3: a superset of the MODE 0 set decided in FP32 in the group-centred frame, the
// radius widened by 1e-6 of the group's half-extent plus k hs_i (sixteen times the FP32
// rounding of the frame's coordinates) and the square by 1e-5 (mixed h-rho pass: the extra
// candidates lie beyond k h_i and carry zero weight).  PER: periodic nearest images.
// MASKS = false (FP32 modes): the walk records the candidates only, and the pair pass that
// reads them makes the same tests on its own staged records.  UNION (MODE 3, no masks): the
// gather walk's candidates also include every particle whose own search radius reaches the
// group box (hcand = the search radii hs_j = 1.05 h_j; a node's hs bound is 1.05 times its h
// maximum), so that while no smoothing length outgrows its search radius the list holds the
// scatter-gather set of the converged smoothing lengths as well.
template <int MODE, bool PER, bool MASKS = true, bool UNION = false>
__global__ __launch_bounds__(64) void k_walk_cand(CandWalk a) {
  constexpr bool F32 = MODE >= 2;      // group-frame FP32 tests
  constexpr bool GATHER = MODE == 0 || MODE == 3;
  __shared__ double4 sc[F32 ? 1 : kWave];   // MODE 0/1: (x, y, z, (k h_j)^2)
  __shared__ float4 fc[F32 ? kWave : 1];    // MODE 2/3: group-frame (x, y, z, (k h_j)^2)
  __shared__ int lstk[kLdsStack];    // node stack; deeper entries go to the global region
  const int lane = threadIdx.x;
  int* stk = a.stacks + (size_t) blockIdx.x*kStackCap;
  const int N = a.N;
  const double krsqd = a.p.kern.rangesqd, kr = a.p.kern.range;
  PhysParams pp = a.p;
  pp.periodic = PER ? 1 : 0;
  GroupQueue q = group_queue(a.glist ? a.glist[a.ncap] : a.ngroups);
  for (;;) {
    const int k = group_next(q, a.counter);
    if (k >= q.nq) break;
    const int g = a.glist ? a.glist[k] : k;
    const int G = a.grp_node[g], gfirst = a.grp_first[g], gcnt = a.grp_count[g];
    const int i = gfirst + lane;
    bool has = lane < gcnt;
    if (has && a.status) has = a.status[i] == 1;
    if (__ballot(has) == 0) {        // nothing left to do in this group
      if (lane == 0) a.ncand[g] = 0;
      continue;
    }
    double xi = 0.0, yi = 0.0, zi = 0.0, hi = 0.0, r2i = 0.0;
    if (has) {
      xi = a.x[i]; yi = a.y[i]; zi = a.z[i]; hi = a.htarget[i];
      r2i = range2_rn(krsqd, hi);
    }
    const double hG = wave_max_d(has ? hi : 0.0);
    // gc, gh: the group node's box, the frame of the FP32 tests.  tc, th: the box the walk
    // tests nodes and members against, that of the group's targets alone (the refit's
    // expressions, so it is the node's box when every member is a target); a block step's
    // sparse groups hold a few targets each and reach far fewer buckets with it
    const double4 gc = a.geo_c[G], gh = a.geo_h[G];
    double4 tc, th;
    {
#pragma clang fp contract(off)
      const double lox = -wave_max_d(has ? -xi : -HUGE_VAL), hix = wave_max_d(has ? xi : -HUGE_VAL);
      const double loy = -wave_max_d(has ? -yi : -HUGE_VAL), hiy = wave_max_d(has ? yi : -HUGE_VAL);
      const double loz = -wave_max_d(has ? -zi : -HUGE_VAL), hiz = wave_max_d(has ? zi : -HUGE_VAL);
      tc = make_double4(0.5*(lox + hix), 0.5*(loy + hiy), 0.5*(loz + hiz), 0.0);
      th = make_double4(0.5*(hix - lox), 0.5*(hiy - loy), 0.5*(hiz - loz), 0.0);
    }
    const float rxi = (float) (xi - gc.x), ryi = (float) (yi - gc.y), rzi = (float) (zi - gc.z);
    const float r2if = (float) r2i;
    float r2gf = 0.f;
    if (MODE == 3 && has) r2gf = gather_r2f(r2i, gh);
    int* idx = a.idx + (size_t) g*a.capb*kWave;
    unsigned long long* msk = a.mask + (size_t) g*a.capb*kWave;
    const int cap = a.capb*kWave;
    int sp = 1, tail = 0;
    bool overflow = false;
    if (lane == 0) lstk[0] = a.grp_root ? a.grp_root[g] : 0;   // the group's segment subtree
    __syncthreads();
    // 1. traversal: the particles of every bucket within range, in walk order, into idx
    while (sp > 0) {
      const int b = min(sp, kWave);
      int node = -1;
      if (lane < b) node = stack_get(lstk, stk, sp - 1 - lane);
      sp -= b;
      bool open = false, emit = false;
      int ef = 0, ec = 0, cl = 0, cr = 0;
      if (node >= 0) {
        // box, range and children in one round of loads
        const double4 nc = a.geo_c[node], nh = a.geo_h[node];
        NodeLinks nl;
        load_links(node, N, a.lnk, nl);
        double dx = nc.x - tc.x, dy = nc.y - tc.y, dz = nc.z - tc.z;
        nearest_image<double>(pp, dx, dy, dz);
        const double gx = fmax(fabs(dx) - nh.x - th.x, 0.0);
        const double gy = fmax(fabs(dy) - nh.y - th.y, 0.0);
        const double gz = fmax(fabs(dz) - nh.z - th.z, 0.0);
        const double dbox2 = gx*gx + gy*gy + gz*gz;
        const double rr = UNION ? kr*fmax(hG, 1.05*nc.w) : GATHER ? kr*hG : kr*fmax(hG, nc.w);
        if (dbox2 <= rr*rr*NODE_SLACK) {
          ef = nl.first; ec = nl.count;
          if (ec <= a.nleaf) emit = true;
          else { open = true; cl = nl.left; cr = nl.right; }
        }
      }
      const unsigned long long om = __ballot(open);
      const int nopen = __popcll(om);
      if (sp + 2*nopen > kLdsStack + kStackCap) { overflow = true; break; }
      if (open) {
        const int o = sp + 2*lanes_below(om);
        stack_put(lstk, stk, o, cr);
        stack_put(lstk, stk, o + 1, cl);
      }
      sp += 2*nopen;
      if (sp > kLdsStack) __threadfence_block();   // entries in the global overflow region
      // a bucket's members that lie beyond the largest range from the group box fail every
      // target's test below: they are dropped here, once per member, instead of being tested
      // on all 64 lanes and staged by every pair pass (about a quarter of the members of the
      // buckets a group reaches).  The round's members are dealt one per lane, so that their
      // loads go out together, and the kept ones are recorded in bucket order.
      const int mcount = emit ? ec : 0;
      const int mincl = wave_incl_scan(mcount);
      const int mtot = __shfl(mincl, 63, 64);
      for (int base = 0; base < mtot; base += kWave) {
        const int s = base + lane;
        // the lane whose bucket holds member s: the first with an inclusive count above s
        int L = 0;
        #pragma unroll
        for (int step = 32; step >= 1; step >>= 1)
          if (__shfl(mincl, L + step - 1, 64) <= s) L += step;
        const int fL = __shfl(ef, L, 64), xL = __shfl(mincl - mcount, L, 64);
        bool kp = false;
        const int j = fL + (s - xL);
        if (s < mtot) {
          double dx = a.x[j] - tc.x, dy = a.y[j] - tc.y, dz = a.z[j] - tc.z;
          nearest_image<double>(pp, dx, dy, dz);
          const double ex = fmax(fabs(dx) - th.x, 0.0);
          const double ey = fmax(fabs(dy) - th.y, 0.0);
          const double ez = fmax(fabs(dz) - th.z, 0.0);
          const double rr = (GATHER && !UNION) ? kr*hG : kr*fmax(hG, a.hcand[j]);
          kp = ex*ex + ey*ey + ez*ez <= rr*rr*KEEP_SLACK;
        }
        const unsigned long long km = __ballot(kp);
        if (kp) {
          const int slot = tail + lanes_below(km);
          if (slot < cap) idx[slot] = j;
        }
        tail += __popcll(km);
      }
    }
    if (overflow) {
      if (lane == 0) atomicOr(&a.flags[2], 1);
      continue;
    }
    const int nbat = (tail + kWave - 1)/kWave;
    if (tail > cap) {                // candidates beyond the set's capacity: grow and repeat
      if (lane == 0) { a.ncand[g] = tail; atomicMax(&a.flags[1], nbat); }
      continue;
    }
    if (!MASKS) {
      if (lane == 0) a.ncand[g] = tail;
      continue;
    }
    __threadfence_block();
    // 2. masks: batch b holds the candidates at walk positions b, b + nbat, b + 2 nbat, ...,
    // so every batch samples the whole candidate region and each target's neighbours spread
    // evenly over the batches (fewer idle lanes in the per-target mask loops of the pair passes)
    // The candidate records of batch b + 1 are loaded while batch b is tested, and the
    // candidate indices of batch b + 2 with them.
    auto nbat_of = [&](int bb) { return (tail - bb + nbat - 1)/nbat; };
    auto record = [&](int j) {
      return make_double4(a.x[j], a.y[j], a.z[j], GATHER ? 0.0 : a.hcand[j]);
    };
    double4 cur = make_double4(0.0, 0.0, 0.0, 0.0), nxt = cur;
    int jn = 0;
    if (lane < nbat_of(0)) cur = record(idx[lane*nbat]);
    if (nbat > 1 && lane < nbat_of(1)) jn = idx[1 + lane*nbat];
    for (int b = 0; b < nbat; b++) {
      const int nb = nbat_of(b);
      if (lane < nb) {
        const double r2j = GATHER ? 0.0 : range2_rn(krsqd, cur.w);
        if (F32) {
          double dx = cur.x - gc.x, dy = cur.y - gc.y, dz = cur.z - gc.z;
          nearest_image<double>(pp, dx, dy, dz);
          fc[lane] = make_float4((float) dx, (float) dy, (float) dz, (float) r2j);
        } else {
          sc[lane] = make_double4(cur.x, cur.y, cur.z, r2j);
        }
      }
      int jn2 = 0;
      if (b + 1 < nbat && lane < nbat_of(b + 1)) nxt = record(jn);
      if (b + 2 < nbat && lane < nbat_of(b + 2)) jn2 = idx[b + 2 + lane*nbat];
      __syncthreads();
      unsigned long long m = 0;
      if (has) {
        // all 64 slots in a fixed-count unrolled loop, so several LDS reads are in flight (a
        // loop over the nb filled slots waited on one read per test); slots past nb are
        // masked off after the loop
        #pragma unroll 16
        for (int q = 0; q < kWave; q++) {
          bool in;
          if (MODE == 3) {
            const float4 c = fc[q];
            in = dot3_f(c.x - rxi, c.y - ryi, c.z - rzi) <= r2gf;
          } else if (MODE == 2) {
            const float4 c = fc[q];
            const float d2 = dot3_f(c.x - rxi, c.y - ryi, c.z - rzi);
            in = (d2 < r2if) | (d2 < c.w);
          } else {
            const double4 c = sc[q];
            double dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
            nearest_image<double>(pp, dx, dy, dz);
            const double d2 = dot3_rn(dx, dy, dz);
            if (MODE == 0) in = d2 <= r2i;
            else in = (d2 < r2i) | (d2 < c.w);
          }
          m |= (unsigned long long) in << q;
        }
        if (nb < kWave) m &= (1ULL << nb) - 1ULL;
      }
      msk[b*kWave + lane] = m;
      __syncthreads();
      cur = nxt;
      jn = jn2;
    }
    if (lane == 0) a.ncand[g] = tail;
  }
}

struct DensArgs {
  int N, ngroups, capb, nopbc;
  const int* grp_node; const int* grp_first; const int* grp_count; const double4* geo_c;
  const double4* geo_h;
  const int* idx; unsigned long long* mask; const int* ncand;
  const double4* posm;
  double* h; double* hsearch; double* rho; double* invomega; double* zeta; double* hfactor;
  double* u; double* sound; double* pressure;
  int* status; int* flags; long long* icount;
  // stars of a gas + star run for their zeta terms (nstar = 0: none, or the terms are off)
  int nstar, star_soft;
  const double* sx; const double* sy; const double* sz; const double* sm; const double* sh;
  PhysParams p;
};

// One step of the grad-h iteration after the sums at h (GradhSph.cpp:219-256): the fixed point
// h <- h_fac (m/rho)^(1/3), bisection on [h_lower, h_upper] after hrho_itmax (30) iterations,
// failure after five times as many.  Returns 2 to iterate again, 1 done (converged, or the bracket closed), 0 retry with a
// larger search radius, -1 no convergence.
__device__ __forceinline__ int hrho_step(const PhysParams& p, double m, double hmax, int iteration,
                                         double rho, double& h, double& h_lower,
                                         double& h_upper) {
  const double invh = 1.0/h;
  const double hrho = rho > 0.0 ? p.h_fac*pow(m/rho, 1.0/3) : 0.0;
  if (rho > 0.0 && h > h_lower && fabs(h - hrho)*invh < p.h_converge) return 1;
  if (iteration < p.hrho_itmax) h = hrho;
  else if (iteration == p.hrho_itmax) h = 0.5*(h_lower + h_upper);
  else if (iteration < 5*p.hrho_itmax) {
    if (rho < 1.0e-20 || h > hrho) h_upper = h; else h_lower = h;
    h = 0.5*(h_lower + h_upper);
  } else return -1;
  if (h > hmax) return 0;
  return (h > h_lower && h < h_upper) ? 2 : 1;
}

// Final h from rho, grad-h terms and EOS (GradhSph.cpp:260-317)
__device__ void hrho_finish(const DensArgs& a, const PhysParams& p, int i, double m, double rho,
                            double invomega, double zeta, double h_lower) {
  const double hrho = p.h_fac*pow(m/rho, 1.0/3);
  const double h = fmax(hrho, h_lower);
  const double invh = 1/h;
  const double hfac4 = invh*invh*invh*invh;
  const double deriv = -(1.0/3)*h/rho;
  double iom = 1.0 - deriv*invomega;
  iom = 1.0/iom;
  // the stars' conservative-gravity zeta terms (Hubber et al. 2013) at the final h, over the
  // pair's mean smoothing length or the particle's own (GradhSph.cpp:282-303)
  if (a.nstar > 0) {
#pragma clang fp contract(off)
    const double4 me = a.posm[i];
    for (int j = 0; j < a.nstar; j++) {
      double invhsqd;
      if (a.star_soft) { const double t = 2.0/(h + a.sh[j]); invhsqd = t*t; }
      else invhsqd = 4.0*invh*invh;
      const double dx = a.sx[j] - me.x, dy = a.sy[j] - me.y, dz = a.sz[j] - me.z;
      const double ssqd = (dx*dx + dy*dy + dz*dz)*invhsqd;
      zeta += a.sm[j]*invhsqd*kwzeta_s2(p.kern, ssqd);
    }
  }
  const double zt = deriv*zeta*iom;
  a.h[i] = h; a.rho[i] = rho; a.invomega[i] = iom; a.zeta[i] = zt; a.hfactor[i] = hfac4;
  double u = a.u[i], snd, press;
  eos_state(p, rho, u, snd, press);
  a.u[i] = u; a.sound[i] = snd; a.pressure[i] = press;
}

// h-rho pass over one walk group per workgroup.  T = double: FP64 sums in candidate order;
// T = float: FP32 pair terms in the group-centred frame, per-batch register sums and
// Kahan-compensated totals (the convergence test and the final h stay FP64).  FUSED (FP32): the
// candidate walk recorded no masks; the first sweep tests each staged batch against the
// target's widened radius exactly as the walk's MODE 3 would, keeps the masks for the later
// sweeps and sums the same pairs.
template <typename T, bool FUSED = false>
__global__ __launch_bounds__(64) void k_density(DensArgs a) {
  __shared__ double4 sd[kWave];
  __shared__ float4 sf[kWave];
  const int g = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int nc = a.ncand[g];
  if (nc <= 0 || nc > a.capb*kWave) return;   // no active target, or the walk overflowed
  const int gfirst = a.grp_first[g], gcnt = a.grp_count[g];
  const int i = gfirst + lane;
  const bool has = lane < gcnt;
  int state = (has && a.status[i] == 1) ? 2 : 3;     // 2 iterating, 3 not a target here
  PhysParams p = a.p;
  if (a.nopbc) p.periodic = 0;
  const double4 gc = a.geo_c[a.grp_node[g]];
  const double4 me = has ? a.posm[i] : make_double4(0.0, 0.0, 0.0, 1.0);
  const double m = me.w;
  double h = has ? a.h[i] : 1.0;
  const double hmax = has ? a.hsearch[i] : 1.0;
  double h_lower = 0.0, h_upper = hmax;
  int iteration = 0;
  long long ngather = 0, evals = 0;
  const float rxi = (float) (me.x - gc.x), ryi = (float) (me.y - gc.y), rzi = (float) (me.z - gc.z);
  const double rsq = p.kern.rangesqd;
  const int nbat = (nc + kWave - 1)/kWave;
  const int* idx = a.idx + (size_t) g*a.capb*kWave;
  unsigned long long* msk = a.mask + (size_t) g*a.capb*kWave;
  float r2gf = 0.f;
  if (FUSED && state == 2) r2gf = gather_r2f(range2_rn(rsq, hmax), a.geo_h[a.grp_node[g]]);
  bool first = true;
  while (__ballot(state == 2)) {
    const bool run = state == 2;
    if (run) iteration++;
    const double invh = 1.0/h;
    const double hfac3 = invh*invh*invh;
    const double invhsqd = invh*invh;
    const float invhf = (float) invh, invhsqdf = (float) invhsqd;
    int inside = 0;
    double r_ = 0.0, om_ = 0.0, z_ = 0.0;
    KahanF kr, kom, kz;
    for (int b = 0; b < nbat; b++) {
      const int nb = (nc - b + nbat - 1)/nbat;   // candidates b, b + nbat, ...
      if (lane < nb) {
        const double4 pj = a.posm[idx[b + lane*nbat]];
        if (sizeof(T) == 8) {
          sd[lane] = pj;
        } else {
          double dx = pj.x - gc.x, dy = pj.y - gc.y, dz = pj.z - gc.z;
          nearest_image<double>(p, dx, dy, dz);
          sf[lane] = make_float4((float) dx, (float) dy, (float) dz, (float) pj.w);
        }
      }
      __syncthreads();
      if (run) {
        unsigned long long mk;
        if (FUSED && first) {
          mk = 0;
          #pragma unroll 16
          for (int q = 0; q < kWave; q++) {
            const float4 c = sf[q];
            mk |= (unsigned long long) (dot3_f(c.x - rxi, c.y - ryi, c.z - rzi) <= r2gf) << q;
          }
          if (nb < kWave) mk &= (1ULL << nb) - 1ULL;
          msk[b*kWave + lane] = mk;
        } else {
          mk = msk[b*kWave + lane];
        }
        evals += __popcll(mk);
        if (sizeof(T) == 8) {
          while (mk) {
            const int q = __builtin_ctzll(mk);
            mk &= mk - 1;
            const double4 pj = sd[q];
            double dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
            nearest_image<double>(p, dx, dy, dz);
            const double ssqd = invhsqd*(dx*dx + dy*dy + dz*dz);
            inside += ssqd < rsq;
            double w0, wom, wz;
            kdens<double>(p.kern, ssqd, w0, wom, wz);
            r_ += pj.w*w0;
            om_ += pj.w*invh*wom;
            z_ += pj.w*wz;
          }
        } else {
          float br = 0.f, bom = 0.f, bz = 0.f;
          while (mk) {
            const int q = __builtin_ctzll(mk);
            mk &= mk - 1;
            const float4 c = sf[q];
            const float dx = c.x - rxi, dy = c.y - ryi, dz = c.z - rzi;
            const float ssqd = invhsqdf*(dx*dx + dy*dy + dz*dz);
            inside += ssqd < (float) rsq;
            float w0, wom, wz;
            kdens<float>(p.kern, ssqd, w0, wom, wz);
            br += c.w*w0;
            bom += c.w*invhf*wom;
            bz += c.w*wz;
          }
          kr.add(br); kom.add(bom); kz.add(bz);
        }
      }
      __syncthreads();
    }
    first = false;
    if (run) {
      double rho, invomega, zeta;
      if (sizeof(T) == 8) { rho = r_; invomega = om_; zeta = z_; }
      else { rho = (double) kr.s; invomega = (double) kom.s; zeta = (double) kz.s; }
      rho *= hfac3;
      invomega *= hfac3;
      zeta *= invhsqd;
      const int rc = hrho_step(p, m, hmax, iteration, rho, h, h_lower, h_upper);
      if (rc == 1) {
        // converged: the sums and the bracket go to k_hrho_finish (status 2), which keeps
        // the final h, grad-h terms and EOS out of this kernel's register budget
        a.rho[i] = rho; a.invomega[i] = invomega; a.zeta[i] = zeta; a.hfactor[i] = h_lower;
        a.status[i] = 2;
        ngather = inside;
        state = 1;
      } else if (rc == 0) {
        a.hsearch[i] = hmax*1.05;       // GradhSphTree.cpp:154: hmax = 1.05 hmax per retry
        atomicAdd(&a.flags[0], 1);
        state = 0;
      } else if (rc < 0) {
        a.status[i] = 0;
        atomicOr(&a.flags[3], 1);
        state = -1;
      }
    }
  }
  wave_add_ll(&a.icount[0], ngather);
  wave_add_ll(&a.icount[4], evals);
}

// Final h, grad-h terms and EOS of the particles the h-rho pass has just converged (status 2:
// rho, invomega and zeta hold the sums at the converged h, hfactor the bracket's lower end)
__global__ void k_hrho_finish(DensArgs a) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= a.N || a.status[i] != 2) return;
  hrho_finish(a, a.p, i, a.posm[i].w, a.rho[i], a.invomega[i], a.zeta[i], a.hfactor[i]);
  a.status[i] = 0;
  // a final h past the search radius (h from rho after the convergence test can exceed it by
  // up to h_converge): the walk's union lists no longer bound the scatter-gather set
  if (a.h[i] > a.hsearch[i]) atomicOr(&a.flags[3], 2);
}

// act (block timesteps): only the active particles iterate; NULL = all
__global__ void k_init_density(int N, const double* __restrict__ h, double* __restrict__ hs,
                               int* __restrict__ status, const int* __restrict__ act) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  hs[i] = 1.05*h[i];      // the search radius and bisection bracket, as GradhSphTree.cpp:154
  status[i] = act ? act[i] : 1;
}

__global__ void k_zero_ints(int n, int* __restrict__ p) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < n) p[i] = 0;
}

// hydro neighbour records: A = (v, 1/h), B = (1/rho, P invomega/rho^2, sound, hfactor)
__global__ void k_pack_hydro(int N, const double* __restrict__ vx, const double* __restrict__ vy,
                             const double* __restrict__ vz, const double* __restrict__ h,
                             const double* __restrict__ rho, const double* __restrict__ pres,
                             const double* __restrict__ iom, const double* __restrict__ snd,
                             const double* __restrict__ hf, double4* __restrict__ A,
                             double4* __restrict__ B) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  A[i] = make_double4(vx[i], vy[i], vz[i], 1/h[i]);
  const double r = rho[i];
  B[i] = make_double4(1/r, (pres[i]*iom[i])/(r*r), snd[i], hf[i]);
}

struct HydroArgs {
  int N, ngroups, capb;
  const int* grp_node; const int* grp_first; const int* grp_count; const double4* geo_c;
  const int* idx; const unsigned long long* mask; const int* ncand;
  const double4* posm; const double4* A; const double4* B;
  const double* h;                  // smoothing lengths, for the fused scatter-gather tests
  const double* invomega; const double* pressure; const double* zeta;
  double* ax; double* ay; double* az; double* dudt; double* div_v;
  double* gx; double* gy; double* gz; double* gpot;    // softened neighbour gravity (VAR 1)
  long long* icount;
  // block timesteps (BLK): active flags, levels and neighbour levels
  const int* act; const int* level; int* levelneib;
  PhysParams p;
};

// Neighbour levels of a block step (GradhSph.cpp:445-446, merged as GradhSphTree.cpp:375-417):
// an active i takes the largest level of its scatter-gather set (itself included), and every
// inactive j of that set takes max(levelneib_j, level_i).  The set relation is symmetric, so an
// active j gets level_i from its own pass; records carry level_j and the bound below which an
// update of j is needed (255 for an active j: never).
__device__ __forceinline__ int blk_record(const HydroArgs& a, int j) {
  return a.level[j] | ((a.act[j] ? 255 : a.levelneib[j]) << 8);
}

// VAR 0: ComputeSphHydroForces (GradhSph.cpp:361-460); VAR 1: ComputeSphHydroGravForces
// (:474-585), the hydro terms plus the kernel-softened gravity of the pair (into gx, gy, gz,
// gpot; the tree walk adds the direct and multipole terms afterwards).  FP64, one walk group
// per workgroup, neighbour records staged in LDS per batch of 64 candidates.
template <int VAR, bool BLK>
__global__ __launch_bounds__(64) void k_hydro(HydroArgs a) {
  __shared__ double4 sP[kWave], sA[kWave], sB[kWave];
  __shared__ double sZ[kWave];
  const int g = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int nc = a.ncand[g];
  const int gfirst = a.grp_first[g], gcnt = a.grp_count[g];
  const int i = gfirst + lane;
  const bool has = lane < gcnt && (!BLK || a.act[gfirst + lane]);
  __shared__ int sJ[kWave], sL[kWave];
  int lmax = 0;
  const int level_i = (BLK && has) ? a.level[i] : 0;
  const PhysParams& p = a.p;
  double4 me = make_double4(0.0, 0.0, 0.0, 1.0), mA = make_double4(0.0, 0.0, 0.0, 1.0);
  double4 mB = make_double4(1.0, 0.0, 0.0, 0.0);
  double zeta_i = 0.0;
  if (has) {
    me = a.posm[i]; mA = a.A[i]; mB = a.B[i];
    if (VAR == 1) zeta_i = a.zeta[i];
  }
  const double invh_i = mA.w, invrho_i = mB.x;
  const double pterm_i = mB.y, snd_i = mB.z, hf_i = mB.w;
  const double alpha = p.alpha_visc, beta = p.beta_visc;
  const double invhsqd_i = invh_i*invh_i;
  double ax = 0, ay = 0, az = 0, du = 0, dv = 0;
  double gx = 0, gy = 0, gz = 0, gp = 0;
  long long npairs = 0;
  const int nbat = (nc + kWave - 1)/kWave;
  const int* idx = a.idx + (size_t) g*a.capb*kWave;
  const unsigned long long* msk = a.mask + (size_t) g*a.capb*kWave;
  for (int b = 0; b < nbat; b++) {
    const int nb = (nc - b + nbat - 1)/nbat;   // candidates b, b + nbat, ...
    if (lane < nb) {
      const int j = idx[b + lane*nbat];
      sP[lane] = a.posm[j]; sA[lane] = a.A[j]; sB[lane] = a.B[j];
      if (VAR == 1) sZ[lane] = a.zeta[j];
      if (BLK) { sJ[lane] = j; sL[lane] = blk_record(a, j); }
    }
    __syncthreads();
    if (has) {
      unsigned long long mk = msk[b*kWave + lane];
      npairs += __popcll(mk);
      while (mk) {
        const int q = __builtin_ctzll(mk);
        mk &= mk - 1;
        if (BLK) {
          const int lj = sL[q];
          lmax = max(lmax, lj & 255);
          if (level_i > (lj >> 8)) atomicMax(&a.levelneib[sJ[q]], level_i);
        }
        const double4 pj = sP[q], Aj = sA[q], Bj = sB[q];
        double dx = pj.x - me.x, dy = pj.y - me.y, dz = pj.z - me.z;
        nearest_image<double>(p, dx, dy, dz);
        const double mj = pj.w;
        const double invh_j = Aj.w, invrho_j = Bj.x;
        double drmag, dvdr;
        if (VAR == 0) {
          drmag = sqrt(dx*dx + dy*dy + dz*dz);
          if (drmag > 0.0) { dx /= drmag; dy /= drmag; dz /= drmag; }
        } else {
          drmag = sqrt(dx*dx + dy*dy + dz*dz + 1.0e-20);
          const double inv = 1.0/drmag;
          dx *= inv; dy *= inv; dz *= inv;
        }
        const double si = drmag*invh_i, sj = drmag*invh_j;
        const double wkerni = hf_i*kw1<double>(p.kern, si);
        const double wkernj = Bj.w*kw1<double>(p.kern, sj);
        if (VAR == 0) {
          dvdr = Aj.x*dx + Aj.y*dy + Aj.z*dz;
          dvdr -= mA.x*dx + mA.y*dy + mA.z*dz;
          dv -= mj*dvdr*wkerni;
        } else {
          const double dvx = Aj.x - mA.x, dvy = Aj.y - mA.y, dvz = Aj.z - mA.z;
          dvdr = dvx*dx + dvy*dy + dvz*dz;
        }
        double paux = pterm_i*wkerni + Bj.y*wkernj;
        if (dvdr < 0.0) {
          const double winvrho = 0.25*(wkerni + wkernj)*(invrho_i + invrho_j);
          const double vsignal = snd_i + Bj.z - beta*alpha*dvdr;
          paux -= alpha*vsignal*dvdr*winvrho;
          du -= 0.5*mj*alpha*vsignal*dvdr*dvdr*winvrho;
        }
        ax += mj*dx*paux; ay += mj*dy*paux; az += mj*dz*paux;
        if (VAR == 1) {
          const double pg = 0.5*(invhsqd_i*kwgrav<double>(p.kern, si) + zeta_i*wkerni +
                                 invh_j*invh_j*kwgrav<double>(p.kern, sj) + sZ[q]*wkernj);
          gx += mj*dx*pg; gy += mj*dy*pg; gz += mj*dz*pg;
          gp += 0.5*mj*(invh_i*kwpot<double>(p.kern, si) + invh_j*kwpot<double>(p.kern, sj));
          dv -= mj*dvdr*wkerni;
        }
      }
    }
    __syncthreads();
  }
  wave_add_ll(&a.icount[1], npairs);
  if (!has) return;
  if (BLK) a.levelneib[i] = lmax;
  const double dvd = dv*mB.x;
  const double dud = du - a.pressure[i]*dvd*mB.x*a.invomega[i];
  a.ax[i] = ax; a.ay[i] = ay; a.az[i] = az;
  a.dudt[i] = dud; a.div_v[i] = dvd;
  if (VAR == 1) { a.gx[i] = gx; a.gy[i] = gy; a.gz[i] = gz; a.gpot[i] = gp; }
}

// FP32 variant: records staged in LDS in the group frame (positions relative to the group
// centre, velocities relative to the group's first particle), pair terms in FP32 with the
// hardware reciprocal and square root, per-batch register sums, Kahan-compensated totals.
// (at most 96 VGPRs: five wavefronts per SIMD instead of four, a few registers spilled)
// FUSED: the candidate walk recorded no masks; each staged batch is tested against both
// smoothing ranges exactly as the walk's MODE 2 would, giving the same scatter-gather set.
template <int VAR, bool BLK, bool FUSED = false>
__global__ __launch_bounds__(64, 5) void k_hydro_mixed(HydroArgs a) {
  __shared__ float4 fP[kWave], fV[kWave], fB[kWave];
  __shared__ float fZ[kWave], fR[kWave];
  const int g = xcd_block(blockIdx.x, gridDim.x);
  const int lane = threadIdx.x;
  const int nc = a.ncand[g];
  const int gfirst = a.grp_first[g], gcnt = a.grp_count[g];
  const int i = gfirst + lane;
  const bool has = lane < gcnt && (!BLK || a.act[gfirst + lane]);
  __shared__ int sJ[kWave], sL[kWave];
  int lmax = 0;
  const int level_i = (BLK && has) ? a.level[i] : 0;
  const PhysParams& p = a.p;
  const double4 gc = a.geo_c[a.grp_node[g]];
  const double4 vref = a.A[gfirst];
  double4 me = make_double4(0.0, 0.0, 0.0, 1.0), mA = make_double4(0.0, 0.0, 0.0, 1.0);
  double4 mB = make_double4(1.0, 0.0, 0.0, 0.0);
  float zeta_i = 0.f;
  if (has) {
    me = a.posm[i]; mA = a.A[i]; mB = a.B[i];
    if (VAR == 1) zeta_i = (float) a.zeta[i];
  }
  const float rxi = (float) (me.x - gc.x), ryi = (float) (me.y - gc.y), rzi = (float) (me.z - gc.z);
  const float vxi = (float) (mA.x - vref.x), vyi = (float) (mA.y - vref.y), vzi = (float) (mA.z - vref.z);
  const float invh_i = (float) mA.w;
  const float invrho_i = (float) mB.x;
  const float pterm_i = (float) mB.y, snd_i = (float) mB.z, hf_i = (float) mB.w;
  const float alpha = (float) p.alpha_visc, beta = (float) p.beta_visc;
  const float invhsqd_i = invh_i*invh_i;
  const double krsqd = p.kern.rangesqd;
  const float r2if = (FUSED && has) ? (float) range2_rn(krsqd, a.h[i]) : 0.f;
  KahanF ax, ay, az, du, dv, gx, gy, gz, gp;
  long long npairs = 0, ntests = 0;
  const int nbat = (nc + kWave - 1)/kWave;
  const int* idx = a.idx + (size_t) g*a.capb*kWave;
  const unsigned long long* msk = a.mask + (size_t) g*a.capb*kWave;
  for (int b = 0; b < nbat; b++) {
    const int nb = (nc - b + nbat - 1)/nbat;   // candidates b, b + nbat, ...
    if (FUSED && has) ntests += nb;
    if (lane < nb) {
      const int j = idx[b + lane*nbat];
      const double4 pj = a.posm[j], Aj = a.A[j], Bj = a.B[j];
      double dx = pj.x - gc.x, dy = pj.y - gc.y, dz = pj.z - gc.z;
      nearest_image<double>(p, dx, dy, dz);
      fP[lane] = make_float4((float) dx, (float) dy, (float) dz, (float) pj.w);
      fV[lane] = make_float4((float) (Aj.x - vref.x), (float) (Aj.y - vref.y),
                             (float) (Aj.z - vref.z), (float) Aj.w);
      fB[lane] = make_float4((float) Bj.x, (float) Bj.y, (float) Bj.z, (float) Bj.w);
      if (VAR == 1) fZ[lane] = (float) a.zeta[j];
      if (FUSED) fR[lane] = (float) range2_rn(krsqd, a.h[j]);
      if (BLK) { sJ[lane] = j; sL[lane] = blk_record(a, j); }
    }
    __syncthreads();
    if (has) {
      unsigned long long mk;
      if (FUSED) {
        mk = 0;
        #pragma unroll 8
        for (int q = 0; q < kWave; q++) {
          const float4 c = fP[q];
          const float d2 = dot3_f(c.x - rxi, c.y - ryi, c.z - rzi);
          mk |= (unsigned long long) ((d2 < r2if) | (d2 < fR[q])) << q;
        }
        if (nb < kWave) mk &= (1ULL << nb) - 1ULL;
      } else {
        mk = msk[b*kWave + lane];
      }
      npairs += __popcll(mk);
      float bax = 0.f, bay = 0.f, baz = 0.f, bdu = 0.f, bdv = 0.f;
      float bgx = 0.f, bgy = 0.f, bgz = 0.f, bgp = 0.f;
      while (mk) {
        const int q = __builtin_ctzll(mk);
        mk &= mk - 1;
        if (BLK) {
          const int lj = sL[q];
          lmax = max(lmax, lj & 255);
          if (level_i > (lj >> 8)) atomicMax(&a.levelneib[sJ[q]], level_i);
        }
        const float4 P = fP[q], V = fV[q], B = fB[q];
        float dx = P.x - rxi, dy = P.y - ryi, dz = P.z - rzi;
        const float mj = P.w;
        const float invh_j = V.w;
        const float invrho_j = B.x;
        float drmag;
        if (VAR == 0) {
          drmag = sqrt_f(dx*dx + dy*dy + dz*dz);
          if (drmag > 0.f) { const float inv = rcp_f(drmag); dx *= inv; dy *= inv; dz *= inv; }
        } else {
          drmag = sqrt_f(dx*dx + dy*dy + dz*dz + 1.0e-20f);
          const float inv = rcp_f(drmag);
          dx *= inv; dy *= inv; dz *= inv;
        }
        const float si = drmag*invh_i, sj = drmag*invh_j;
        const float wkerni = hf_i*kw1<float>(p.kern, si);
        const float wkernj = B.w*kw1<float>(p.kern, sj);
        const float dvx = V.x - vxi, dvy = V.y - vyi, dvz = V.z - vzi;
        const float dvdr = dvx*dx + dvy*dy + dvz*dz;
        bdv -= mj*dvdr*wkerni;
        float paux = pterm_i*wkerni + B.y*wkernj;
        if (dvdr < 0.f) {
          const float winvrho = 0.25f*(wkerni + wkernj)*(invrho_i + invrho_j);
          const float vsignal = snd_i + B.z - beta*alpha*dvdr;
          paux -= alpha*vsignal*dvdr*winvrho;
          bdu -= 0.5f*mj*alpha*vsignal*dvdr*dvdr*winvrho;
        }
        bax += mj*dx*paux; bay += mj*dy*paux; baz += mj*dz*paux;
        if (VAR == 1) {
          float wgi, wpi, wgj, wpj;
          kgravpot_f(p.kern, si, rcp_f(si), wgi, wpi);
          kgravpot_f(p.kern, sj, rcp_f(sj), wgj, wpj);
          const float pg = 0.5f*(invhsqd_i*wgi + zeta_i*wkerni + invh_j*invh_j*wgj +
                                 fZ[q]*wkernj);
          bgx += mj*dx*pg; bgy += mj*dy*pg; bgz += mj*dz*pg;
          bgp += 0.5f*mj*(invh_i*wpi + invh_j*wpj);
        }
      }
      ax.add(bax); ay.add(bay); az.add(baz); du.add(bdu); dv.add(bdv);
      if (VAR == 1) { gx.add(bgx); gy.add(bgy); gz.add(bgz); gp.add(bgp); }
    }
    __syncthreads();
  }
  wave_add_ll(&a.icount[1], npairs);
  if (FUSED) wave_add_ll(&a.icount[5], ntests);
  if (!has) return;
  if (BLK) a.levelneib[i] = lmax;
  const double dvd = (double) dv.s*mB.x;
  const double dud = (double) du.s - a.pressure[i]*dvd*mB.x*a.invomega[i];
  a.ax[i] = (double) ax.s; a.ay[i] = (double) ay.s; a.az[i] = (double) az.s;
  a.dudt[i] = dud; a.div_v[i] = dvd;
  if (VAR == 1) {
    a.gx[i] = (double) gx.s; a.gy[i] = (double) gy.s; a.gz[i] = (double) gz.s;
    a.gpot[i] = (double) gp.s;
  }
}

int read_flags(gdf_ctx* c) {
  GDF_HIP(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(int)*8, hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

void free_cand(CandSet& cs) {
  if (cs.idx) (void) hipFree(cs.idx);
  if (cs.mask) (void) hipFree(cs.mask);
  if (cs.ncand) (void) hipFree(cs.ncand);
  cs.idx = nullptr; cs.mask = nullptr; cs.ncand = nullptr;
  cs.capb = 0; cs.gcap = 0;
}

// (Re)allocate a candidate set for at least the current number of groups and capb batches.
int alloc_cand(gdf_ctx* c, CandSet& cs, int capb) {
  const int gcap = std::max(cs.gcap, c->ngroups + c->ngroups/4 + 64);
  if (cs.idx && gcap == cs.gcap && capb == cs.capb) return GDF_OK;
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  free_cand(cs);
  const size_t n = (size_t) gcap*(size_t) capb*kWave;
  if (hipMalloc(&cs.idx, sizeof(int)*n) != hipSuccess ||
      hipMalloc(&cs.mask, sizeof(unsigned long long)*n) != hipSuccess ||
      hipMalloc(&cs.ncand, sizeof(int)*(size_t) gcap) != hipSuccess) {
    free_cand(cs);
    return fail(c, GDF_ERR_NOMEM, "candidate set allocation failed");
  }
  cs.gcap = gcap;
  cs.capb = capb;
  return GDF_OK;
}

int ensure_cand(gdf_ctx* c, CandSet& cs) {
  if (cs.idx && c->ngroups <= cs.gcap) return GDF_OK;
  return alloc_cand(c, cs, cs.capb > 0 ? cs.capb
                            : (c->cfg.cand_batches_init > 0 ? c->cfg.cand_batches_init
                                                            : kInitBatches));
}

int grow_cand(gdf_ctx* c, CandSet& cs, int need) {
  return alloc_cand(c, cs, std::max(need + need/4 + 1, cs.capb + 1));
}

// glist: the groups holding targets (from group_list), or NULL to walk every group
// (the walk's queue counters, counters[4, 12), must be zero: the list builder clears them)
// masks = false (mixed path): the FP32 walks record candidates only (the pair pass tests them);
// uni: the mixed gather walk's union lists (gather and scatter-gather candidates)
int launch_walk(gdf_ctx* c, int mode, int nopbc, CandSet& cs, const int* status,
                const int* glist, bool masks, bool uni = false) {
  CandWalk w;
  w.N = c->N; w.ngroups = c->ngroups; w.nleaf = c->phys.nleaf; w.capb = cs.capb;
  w.grp_node = c->grp_node; w.grp_first = c->grp_first; w.grp_count = c->grp_count;
  w.lnk = c->lnk;
  w.geo_c = c->geo_c; w.geo_h = c->geo_h;
  w.x = c->pf[F_X]; w.y = c->pf[F_Y]; w.z = c->pf[F_Z];
  w.hcand = uni ? c->df[D_HSEARCH] : c->pf[F_H];
  w.htarget = mode == 0 ? c->df[D_HSEARCH] : c->pf[F_H];
  w.status = status;
  w.grp_root = c->nseg > 0 ? c->grp_root : nullptr;
  w.glist = glist; w.ncap = c->Ncap;
  w.idx = cs.idx; w.mask = cs.mask; w.ncand = cs.ncand;
  w.stacks = c->stacks; w.counter = c->counters + 4; w.flags = c->counters;
  w.p = c->phys;
  if (nopbc) w.p.periodic = 0;
  if (!glist) k_zero_ints<<<1, BS, 0, c->stream>>>(8, c->counters + 4);
  const int grid = std::min(c->ngroups, kWalkSlots);
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  const int m = mode == 0 ? (mixed ? 3 : 0) : (mixed ? 2 : 1);
  hipStream_t st = c->stream;
  if (w.p.periodic) {
    if (m == 0) k_walk_cand<0, true><<<grid, kWave, 0, st>>>(w);
    else if (m == 1) k_walk_cand<1, true><<<grid, kWave, 0, st>>>(w);
    else if (m == 2) {
      if (masks) k_walk_cand<2, true><<<grid, kWave, 0, st>>>(w);
      else k_walk_cand<2, true, false><<<grid, kWave, 0, st>>>(w);
    } else {
      if (masks) k_walk_cand<3, true><<<grid, kWave, 0, st>>>(w);
      else if (uni) k_walk_cand<3, true, false, true><<<grid, kWave, 0, st>>>(w);
      else k_walk_cand<3, true, false><<<grid, kWave, 0, st>>>(w);
    }
  } else {
    if (m == 0) k_walk_cand<0, false><<<grid, kWave, 0, st>>>(w);
    else if (m == 1) k_walk_cand<1, false><<<grid, kWave, 0, st>>>(w);
    else if (m == 2) {
      if (masks) k_walk_cand<2, false><<<grid, kWave, 0, st>>>(w);
      else k_walk_cand<2, false, false><<<grid, kWave, 0, st>>>(w);
    } else {
      if (masks) k_walk_cand<3, false><<<grid, kWave, 0, st>>>(w);
      else if (uni) k_walk_cand<3, false, false, true><<<grid, kWave, 0, st>>>(w);
      else k_walk_cand<3, false, false><<<grid, kWave, 0, st>>>(w);
    }
  }
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

}  // namespace

int sph_alloc_lists(gdf_ctx* c) {
  GDF_HIP(c, hipMalloc(&c->status, sizeof(int)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->hrecA, sizeof(double4)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->hrecB, sizeof(double4)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->stacks, sizeof(int)*(size_t) kStackSlots*kStackCap));
  GDF_HIP(c, hipMalloc(&c->glist, sizeof(int)*3*((size_t) c->Ncap + 1)));
  if (c->phys.self_gravity && c->phys.hydro_forces && !c->phys.nbody)
    GDF_HIP(c, hipMalloc(&c->stacks2, sizeof(int)*(size_t) kStackSlots*kStackCap));
  return GDF_OK;
}

void sph_free_lists(gdf_ctx* c) {
  free_cand(c->cs_g);
  free_cand(c->cs_s);
  if (c->sg_tgt) (void) hipFree(c->sg_tgt);
  c->sg_tgt = nullptr;
  c->sg_tgt_cap = 0;
  c->sg_tgt_all = true;
}

int sph_density(gdf_ctx* c, int nopbc) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "density pass before tree build");
  GDF_TRY(ensure_cand(c, c->cs_g));
  k_init_density<<<nblocks(N, BS), BS, 0, s>>>(N, c->pf[F_H], c->df[D_HSEARCH], c->status,
                                               c->nlevels > 1 ? c->pi[I_ACT] : nullptr);
  DensArgs d;
  d.N = N; d.ngroups = c->ngroups; d.nopbc = nopbc;
  d.grp_node = c->grp_node; d.grp_first = c->grp_first; d.grp_count = c->grp_count;
  d.geo_c = c->geo_c; d.geo_h = c->geo_h;
  d.posm = c->posm;
  d.h = c->pf[F_H]; d.hsearch = c->df[D_HSEARCH]; d.rho = c->df[D_RHO];
  d.invomega = c->df[D_INVOMEGA]; d.zeta = c->df[D_ZETA]; d.hfactor = c->df[D_HFACTOR];
  d.u = c->pf[F_U]; d.sound = c->df[D_SOUND]; d.pressure = c->df[D_PRESSURE];
  d.status = c->status; d.flags = c->counters; d.icount = c->icount;
  d.nstar = c->cfg.conservative_sph_star_gravity ? c->Nstar : 0;
  d.star_soft = c->cfg.nbody_softening;
  d.sx = c->sv[S_X]; d.sy = c->sv[S_Y]; d.sz = c->sv[S_Z]; d.sm = c->sv[S_M]; d.sh = c->sv[S_H];
  d.p = c->phys;
  c->sg_union = false;
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  // mixed path with periodic images: the first round walks the union lists; they stay valid for
  // the force pass if that round settles every target (a regrowth or retry round would walk
  // again only the unsettled groups, and a retry grows search radii past the lists' bounds)
  const bool uni = mixed && !nopbc;
  for (int round = 0; round < 64; round++) {
    // also clears the round's flags and the walk's queue counters
    GDF_TRY(group_list(c, c->status, c->glist, c->cs_g.ncand, nullptr, nullptr, c->counters, 12));
    GDF_TRY(launch_walk(c, 0, nopbc, c->cs_g, c->status, c->glist, !mixed, uni));
    d.capb = c->cs_g.capb; d.idx = c->cs_g.idx; d.mask = c->cs_g.mask; d.ncand = c->cs_g.ncand;
    if (mixed) k_density<float, true><<<c->ngroups, kWave, 0, s>>>(d);
    else k_density<double><<<c->ngroups, kWave, 0, s>>>(d);
    k_hrho_finish<<<nblocks(N, BS), BS, 0, s>>>(d);
    GDF_HIP(c, hipGetLastError());
    GDF_TRY(read_flags(c));
    const int* f = c->h_counters;
    if (f[2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in gather pass");
    if (f[3] & 1) return fail(c, GDF_ERR_NOCONV, "Problem with convergence of h-rho iteration");
    // groups whose candidates overflowed skipped the density kernel and keep status 1
    if (f[1] > 0) { GDF_TRY(grow_cand(c, c->cs_g, f[1])); continue; }
    if (f[0] == 0) {
      c->sg_valid = false;
      GDF_TRY(tree_refit_h(c));
      c->sg_union = uni && round == 0 && !(f[3] & 2);
      return GDF_OK;
    }
  }
  return fail(c, GDF_ERR_CAPACITY, "h-rho neighbour search did not settle after 64 rounds");
}

int sph_build_sg(gdf_ctx* c, bool masks) {
  // the mixed hydro kernel tests its candidates itself unless the masks are asked for
  masks = masks || c->cfg.precision != GDF_PREC_MIXED;
  c->sg_from_union = false;
  if (!masks && c->sg_union) {      // the density pass's union lists hold the set
    c->sg_from_union = true;
    c->sg_valid = true;
    c->sg_masks = false;
    return GDF_OK;
  }
  GDF_TRY(ensure_cand(c, c->cs_s));
  for (int round = 0; round < 8; round++) {
    const int* act = c->nlevels > 1 ? c->pi[I_ACT] : nullptr;
    int* sl = c->glist + (c->Ncap + 1);
    if (act) GDF_TRY(group_list(c, act, sl, c->cs_s.ncand, nullptr, nullptr, c->counters, 12));
    else k_zero_ints<<<1, BS, 0, c->stream>>>(4, c->counters);
    GDF_TRY(launch_walk(c, 1, 0, c->cs_s, act, act ? sl : nullptr, masks));
    GDF_TRY(read_flags(c));
    if (c->h_counters[2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in force pass");
    if (c->h_counters[1] > 0) { GDF_TRY(grow_cand(c, c->cs_s, c->h_counters[1])); continue; }
    c->sg_valid = true;
    c->sg_masks = masks;
    return GDF_OK;
  }
  return fail(c, GDF_ERR_CAPACITY, "scatter-gather candidate capacity");
}

int sph_hydro(gdf_ctx* c, int variant) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (!c->sg_valid) GDF_TRY(sph_build_sg(c));
  k_pack_hydro<<<nblocks(N, BS), BS, 0, s>>>(N, c->pf[F_VX], c->pf[F_VY], c->pf[F_VZ],
                                             c->pf[F_H], c->df[D_RHO], c->df[D_PRESSURE],
                                             c->df[D_INVOMEGA], c->df[D_SOUND],
                                             c->df[D_HFACTOR], c->hrecA, c->hrecB);
  HydroArgs h;
  const CandSet& cs = c->sg_from_union ? c->cs_g : c->cs_s;
  h.N = N; h.ngroups = c->ngroups; h.capb = cs.capb;
  h.grp_node = c->grp_node; h.grp_first = c->grp_first; h.grp_count = c->grp_count;
  h.geo_c = c->geo_c;
  h.idx = cs.idx; h.mask = cs.mask; h.ncand = cs.ncand;
  h.posm = c->posm; h.A = c->hrecA; h.B = c->hrecB; h.h = c->pf[F_H];
  h.invomega = c->df[D_INVOMEGA]; h.pressure = c->df[D_PRESSURE]; h.zeta = c->df[D_ZETA];
  h.ax = c->df[D_AX]; h.ay = c->df[D_AY]; h.az = c->df[D_AZ];
  h.gx = c->df[D_ATX]; h.gy = c->df[D_ATY]; h.gz = c->df[D_ATZ]; h.gpot = c->df[D_GPOT];
  h.dudt = c->df[D_DUDT]; h.div_v = c->df[D_DIVV]; h.p = c->phys;
  h.icount = c->icount;
  h.act = c->pi[I_ACT]; h.level = c->pi[I_LEVEL]; h.levelneib = c->pi[I_LEVELNEIB];
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  const bool blk = c->nlevels > 1;
  const int G = c->ngroups;
  if (mixed) {
    // the pass's targets, for decoding the sets it evaluates (its lists cover those alone)
    c->sg_tgt_all = !blk;
    if (blk) {
      if (c->sg_tgt_cap < N) {
        if (c->sg_tgt) (void) hipFree(c->sg_tgt);
        c->sg_tgt = nullptr;
        c->sg_tgt_cap = 0;
        GDF_HIP(c, hipMalloc(&c->sg_tgt, sizeof(int)*(size_t) c->Ncap));
        c->sg_tgt_cap = c->Ncap;
      }
      GDF_HIP(c, hipMemcpyAsync(c->sg_tgt, c->pi[I_ACT], sizeof(int)*(size_t) N,
                                hipMemcpyDeviceToDevice, s));
    }
  }
  const int kt = ktime_begin(c, s, 6);
  const bool fused = mixed && !c->sg_masks;
  if (variant == 0) {
    if (blk) {
      if (fused) k_hydro_mixed<0, true, true><<<G, kWave, 0, s>>>(h);
      else if (mixed) k_hydro_mixed<0, true><<<G, kWave, 0, s>>>(h);
      else k_hydro<0, true><<<G, kWave, 0, s>>>(h);
    } else {
      if (fused) k_hydro_mixed<0, false, true><<<G, kWave, 0, s>>>(h);
      else if (mixed) k_hydro_mixed<0, false><<<G, kWave, 0, s>>>(h);
      else k_hydro<0, false><<<G, kWave, 0, s>>>(h);
    }
  } else {
    if (blk) {
      if (fused) k_hydro_mixed<1, true, true><<<G, kWave, 0, s>>>(h);
      else if (mixed) k_hydro_mixed<1, true><<<G, kWave, 0, s>>>(h);
      else k_hydro<1, true><<<G, kWave, 0, s>>>(h);
    } else {
      if (fused) k_hydro_mixed<1, false, true><<<G, kWave, 0, s>>>(h);
      else if (mixed) k_hydro_mixed<1, false><<<G, kWave, 0, s>>>(h);
      else k_hydro<1, false><<<G, kWave, 0, s>>>(h);
    }
  }
  GDF_HIP(c, hipGetLastError());
  ktime_end(c, s, kt);
  return GDF_OK;
}

int sph_copy_sg(gdf_ctx* c, std::vector<int>& idx, std::vector<unsigned long long>& mask,
                std::vector<int>& ncand, int& capb) {
  // the mixed force pass leaves no masks: the set is walked again with them (same tests)
  if (!c->sg_valid || !c->sg_masks) GDF_TRY(sph_build_sg(c, true));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  const CandSet& cs = c->cs_s;
  const size_t n = (size_t) c->ngroups*(size_t) cs.capb*kWave;
  idx.resize(n); mask.resize(n); ncand.resize(c->ngroups);
  capb = cs.capb;
  GDF_HIP(c, hipMemcpy(idx.data(), cs.idx, sizeof(int)*n, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(mask.data(), cs.mask, sizeof(unsigned long long)*n, hipMemcpyDeviceToHost));
  GDF_HIP(c, hipMemcpy(ncand.data(), cs.ncand, sizeof(int)*c->ngroups, hipMemcpyDeviceToHost));
  return GDF_OK;
}

}  // namespace gdf
