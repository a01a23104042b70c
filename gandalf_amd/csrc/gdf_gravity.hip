// Tree-gravity walk with per-wavefront interaction lists.
//
// Replaces the gravity half of GradhSphTree::UpdateAllSphForces
// (src/GradhSph/GradhSphTree.cpp:444-657): the walk of Tree::ComputeGravityInteractionAndGhostList
// (src/Tree/Tree.cpp:628-735), the kernel-softened neighbour gravity of
// GradhSph::ComputeSphHydroGravForces (src/GradhSph/GradhSph.cpp:555-565), the direct sum of
// GradhSphBase::ComputeDirectGravForces (:657-690) and the quadrupole cell terms of
// ComputeQuadropole (src/Headers/NeighbourSearch.h:423-449).
//
// One wavefront per walk group.  Acceptance per node, with the group's bounding box G:
//   * kernel-range test (as the reference's neighbour-cell test, Tree.cpp:662-663): if the
//     minimum box-box distance is below k*max(hmax_G, hmax_cell) (with a 1e-4 relative slack
//     on the squared distance) the node is resolved to particles (bucket of <= Nleafmax
//     particles) or opened;
//   * geometric MAC (Tree.h:413-432) evaluated from the nearest point of G: accept the
//     multipole iff dist(G, centre_cell)^2 >= cdistsqd = max(rmax^2, hmax^2)/theta^2, a
//     conservative form of the reference's centre-to-centre test; the gadget2 and eigenmac
//     criteria then open more cells, from the same distance (mac_opens);
//   * otherwise open, or send the bucket's particles to the direct list.
// As in the reference, the kernel-softened gravity of the scatter-gather neighbours is summed by
// the hydro pass (GradhSph::ComputeSphHydroGravForces); the walk adds the direct 1/r^2 terms of
// every other particle and the multipole terms, masking the neighbour pairs with the same range
// test.  Without a hydro pass the walk evaluates the softened pairs itself.  Accepted cells and
// direct particles are staged 64 at a time in LDS and every lane (one target particle) sweeps
// the batch; the FP32 path sums each batch in registers and carries Kahan-compensated totals
// across batches.
#include <hip/hip_runtime.h>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int PCAP = 1024;   // particle ring buffer (power of two); Nleafmax <= 14
constexpr int CCAP = 256;    // cell ring buffer (power of two)

struct GravArgs {
  int N, ngroups, nleaf;
  const int* grp_node; const int* grp_first; const int* grp_count;
  const int4* lnk;
  const double4* geo_c; const double4* geo_h; const double* mp;
  const double4* posm; const double4* grec;     // grec: (h, zeta, hfactor, 0)
  double* atx; double* aty; double* atz; double* gpot;
  int* stacks; int* counter; int* flags; long long* icount;
  const int* glist; int ncap;   // groups to walk (count at glist[ncap]); NULL = all groups
  int32_t* gnc; int32_t* gnn;
  const int* act;         // block timesteps: only active targets (NULL = all)
  const int* grp_root;    // serving batches: walk from the group's segment subtree (NULL = root)
  const double* gmac;     // per group: factor of the gadget2 / eigenmac criterion
  // non-NULL: the walk runs beside the hydro pass and writes its own sums here (acceleration,
  // potential, self term) instead of adding to the hydro pass's atx/gpot
  double* gwx; double* gwy; double* gwz; double* gwp; double* gwp0;
  PhysParams p;
};

// Unified FP64 pair term for walks that also own the neighbour pairs (no hydro pass): the
// kernel-softened form of GradhSph.cpp:555-565 inside either smoothing range, the direct form
// of GradhSphBase::ComputeDirectGravForces (:657-690) outside
__device__ __forceinline__ void pair_fp64(const KernelDesc& k, double krsqd, double dx, double dy,
                                          double dz, double mj, double hi, double hj, double zi,
                                          double zj, double hfi, double hfj, double& ax,
                                          double& ay, double& az, double& pot) {
  const double d2 = dx*dx + dy*dy + dz*dz;
  if (d2 < krsqd*hi*hi || d2 < krsqd*hj*hj) {
    const double invh_i = 1/hi, invh_j = 1/hj;
    const double drmag = sqrt(d2 + 1.0e-20);
    const double inv = 1.0/drmag;
    dx *= inv; dy *= inv; dz *= inv;
    const double si = drmag*invh_i, sj = drmag*invh_j;
    const double wkerni = hfi*kw1<double>(k, si);
    const double wkernj = hfj*kw1<double>(k, sj);
    const double paux = 0.5*(invh_i*invh_i*kwgrav<double>(k, si) + zi*wkerni +
                             invh_j*invh_j*kwgrav<double>(k, sj) + zj*wkernj);
    ax += mj*dx*paux; ay += mj*dy*paux; az += mj*dz*paux;
    pot += 0.5*mj*(invh_i*kwpot<double>(k, si) + invh_j*kwpot<double>(k, sj));
  } else {
    const double inv = rsq_d(d2 + 1.0e-20);
    const double inv3 = inv*inv*inv;
    ax += mj*dx*inv3; ay += mj*dy*inv3; az += mj*dz*inv3;
    pot += mj*inv;
  }
}

__device__ __forceinline__ void pair_fp32(const KernelDesc& k, float d2, bool soft, float dx,
                                          float dy, float dz, float mj, float hi, float hj,
                                          float zi, float zj, float hfi, float hfj, float& ax,
                                          float& ay, float& az, float& pot) {
  if (soft) {
    const float invh_i = rcp_f(hi), invh_j = rcp_f(hj);
    const float drmag = sqrt_f(d2 + 1.0e-20f);
    const float inv = rcp_f(drmag);
    dx *= inv; dy *= inv; dz *= inv;
    const float si = drmag*invh_i, sj = drmag*invh_j;
    const float wkerni = hfi*kw1<float>(k, si);
    const float wkernj = hfj*kw1<float>(k, sj);
    const float paux = 0.5f*(invh_i*invh_i*kwgrav<float>(k, si) + zi*wkerni +
                             invh_j*invh_j*kwgrav<float>(k, sj) + zj*wkernj);
    ax += mj*dx*paux; ay += mj*dy*paux; az += mj*dz*paux;
    pot += 0.5f*mj*(invh_i*kwpot<float>(k, si) + invh_j*kwpot<float>(k, sj));
  } else {
    const float inv = rsq_f(d2 + 1.0e-20f);
    const float inv3 = inv*inv*inv;
    ax += mj*dx*inv3; ay += mj*dy*inv3; az += mj*dz*inv3;
    pot += mj*inv;
  }
}

// quadrupole cell term with dr = rp - com (NeighbourSearch.h:423-449); 1/|dr| from the
// reciprocal square root instead of a division and a square root
template <typename T>
__device__ __forceinline__ void quad_term(T dx, T dy, T dz, T mc, const T* q, T& ax, T& ay,
                                          T& az, T& pot, T soft2 = T(1.0e-20)) {
  const T drsqd = dx*dx + dy*dy + dz*dz + soft2;
  T invdrmag;
  if (sizeof(T) == 4) invdrmag = (T) rsq_f((float) drsqd);
  else invdrmag = (T) rsq_d((double) drsqd);
  const T invdrsqd = invdrmag*invdrmag;
  const T invdr5 = invdrsqd*invdrsqd*invdrmag;
  const T mono = mc*invdrsqd*invdrmag;
  const T qscalar = q[0]*dx*dx + q[2]*dy*dy - (q[0] + q[2])*dz*dz +
                    T(2)*(q[1]*dx*dy + q[3]*dx*dz + q[4]*dy*dz);
  const T qfactor = T(2.5)*qscalar*invdr5*invdrsqd;
  ax += -mono*dx + (q[0]*dx + q[1]*dy + q[3]*dz)*invdr5 - qfactor*dx;
  ay += -mono*dy + (q[1]*dx + q[2]*dy + q[4]*dz)*invdr5 - qfactor*dy;
  az += -mono*dz + (q[3]*dx + q[4]*dy - (q[0] + q[2])*dz)*invdr5 - qfactor*dz;
  pot += mc*invdrmag + T(0.5)*qscalar*invdr5;
}

// Tree::open_cell_for_gravity (Tree.h:413-432) past its geometric test, for a cell at squared
// distance d2 from the walk group's box: gadget2 opens a cell whose m rmax^2 / d^4 exceeds
// macerror times the group's smallest |atree| (gfac); eigenmac one nearer than
// (0.5 lambda / macerror)^(2/3) (KDTree.cpp:1058-1079, lambda = 2 sqrt(-p/3) bounding the
// quadrupole's largest eigenvalue) times the group's largest gpot^(-2/3) (gfac)
__device__ __forceinline__ bool mac_opens(const PhysParams& p, const double4& nh, const double* mp,
                                          double d2, double gfac) {
#pragma clang fp contract(off)
  if (p.gravity_mac == GDF_MAC_GADGET2) {
    const double rmax = sqrt(nh.x*nh.x + nh.y*nh.y + nh.z*nh.z);
    return d2*d2*gfac*p.macerror < rmax*rmax*mp[3];
  }
  const double* q = mp + 4;
  const double pq = q[0]*q[2] - (q[0] + q[2])*(q[0] + q[2]) - q[1]*q[1] - q[3]*q[3] - q[4]*q[4];
  const double lambda = pq >= 0.0 ? 0.0 : 2.0*sqrt(-pq/3.0);
  const double mac = pow(0.5*lambda/p.macerror, 0.66666666666666);
  return d2 < mac*gfac;
}

// FastMultipoleForces::AddMonopoleContribution / AddQuadrupoleContribution
// (NeighbourSearch.h:564-597, 675-723) with the second-order terms off: one accepted cell's
// potential, acceleration and tidal tensor at the walk group's expansion point, added to
// f = (pot, ax, ay, az, txx, txy, tyy, txz, tyz, tzz).  dr = cell - expansion point; the
// reference's dphi takes the same increments as its ac, so one vector serves both.
__device__ __forceinline__ void fast_add(double* f, double dx, double dy, double dz, double mc,
                                         const double* q, bool quad) {
  {
    const double invdrmag = sqrt(1.0/(dx*dx + dy*dy + dz*dz));
    const double invdrsqd = invdrmag*invdrmag;
    const double mc3 = mc*invdrsqd*invdrmag;
    f[0] += mc*invdrmag;
    f[1] += mc3*dx; f[2] += mc3*dy; f[3] += mc3*dz;
    f[4] += mc3*(3.0*dx*dx*invdrsqd - 1.0);
    f[5] += mc3*(3.0*dx*dy*invdrsqd);
    f[6] += mc3*(3.0*dy*dy*invdrsqd - 1.0);
    f[7] += mc3*(3.0*dz*dx*invdrsqd);
    f[8] += mc3*(3.0*dz*dy*invdrsqd);
    f[9] += mc3*(3.0*dz*dz*invdrsqd - 1.0);
  }
  if (!quad) return;
  // the quadrupole part is written with d = expansion point - cell
  const double ex = -dx, ey = -dy, ez = -dz;
  const double invdrsqd = 1.0/(ex*ex + ey*ey + ez*ez + 1.0e-20);
  const double invdrmag = sqrt(invdrsqd);
  const double invdr5 = invdrsqd*invdrsqd*invdrmag;
  const double qzz = -(q[0] + q[2]);
  const double qscalar = q[0]*ex*ex + q[2]*ey*ey + qzz*ez*ez +
                         2.0*(q[1]*ex*ey + q[3]*ex*ez + q[4]*ey*ez);
  const double qfactor = 2.5*qscalar*invdr5*invdrsqd;
  double qx = (q[0]*ex + q[1]*ey + q[3]*ez)*invdr5;
  double qy = (q[1]*ex + q[2]*ey + q[4]*ez)*invdr5;
  double qz = (q[3]*ex + q[4]*ey + qzz*ez)*invdr5;
  f[0] += 0.5*qscalar*invdr5;
  f[1] += qx - qfactor*ex; f[2] += qy - qfactor*ey; f[3] += qz - qfactor*ez;
  const double s5 = 5.0*invdrsqd;
  qx *= s5; qy *= s5; qz *= s5;
  f[4] += qfactor*(7.0*ex*ex*invdrsqd - 1.0) - (2.0*qx*ex - q[0]*invdr5);
  f[5] += qfactor*(7.0*ex*ey*invdrsqd) - (qx*ey + qy*ex - q[1]*invdr5);
  f[6] += qfactor*(7.0*ey*ey*invdrsqd - 1.0) - (2.0*qy*ey - q[2]*invdr5);
  f[7] += qfactor*(7.0*ex*ez*invdrsqd) - (qx*ez + qz*ex - q[3]*invdr5);
  f[8] += qfactor*(7.0*ey*ez*invdrsqd) - (qy*ez + qz*ey - q[4]*invdr5);
  f[9] += qfactor*(7.0*ez*ez*invdrsqd - 1.0) - (2.0*qz*ez - qzz*invdr5);
}

// SOFT: the walk also evaluates the kernel-softened neighbour pairs (no hydro pass ran).
// Otherwise the hydro pass has already written the softened gravity of every scatter-gather
// pair into atree/gpot (GradhSph::ComputeSphHydroGravForces), the walk masks exactly those pairs
// (same squared distance, same range test, so each pair lands in one of the two sums) and adds
// the direct and multipole terms.  PLUMMER (pure N-body stars): every particle pair and every
// multipole term is softened with the Plummer length eps, the star itself is skipped and
// there is no self potential.  EWALD (periodic box with self-gravity): every particle of the
// walk's buckets (the hydro pass's neighbours and the target itself included) and every accepted
// cell adds its periodic correction at its separation, as GradhSphTree::UpdateAllSphForces does
// for all its neighbours and cells (GradhSphTree.cpp:558-578).  FAST (multipole fast_monopole or
// fast_quadrupole, gas walks): the accepted cells are summed once per group into the Taylor
// coefficients about the group's centre of mass (each lane takes the cells of its batch slot,
// the lanes' sums are reduced at the end of the walk) and every target applies the first-order
// series (ComputeFastMonopoleForces / ComputeFastQuadrupoleForces, NeighbourSearch.h:769-830).
// A cell is then accepted only if its centre of mass also lies beyond the group's extent about
// that point divided by theta, so that the series converges for every target.  multipole
// monopole: the cells' moments are staged as zeros (the quadrupole term reduces to
// ComputeCellMonopoleForces, NeighbourSearch.h:350-377).
template <bool MIXED, bool SOFT, bool PLUMMER, bool EWALD = false, bool FAST = false>
__global__ __launch_bounds__(64) void k_gravity(GravArgs a) {
  __shared__ int pbuf[PCAP];
  __shared__ int cbuf[CCAP];
  __shared__ int sidx[kWave];
  // staged batch: particles (x, y, z, m, r2, h, zeta, hfactor) or cells (com x3, m, q x5).
  // The FP32 path stages per lane two float4 records and one float, each read back with one
  // broadcast LDS load: particles (dx, dy, dz, m), (h, zeta, hfactor, -), r2; cells
  // (dx, dy, dz, m), (q0, q1, q2, q3), q4
  __shared__ __align__(16) double stage[(MIXED ? 5 : 9)*kWave];
  double (*sd)[kWave] = reinterpret_cast<double (*)[kWave]>(stage);
  float4* s4a = reinterpret_cast<float4*>(stage);
  float4* s4b = s4a + kWave;
  float* s1 = reinterpret_cast<float*>(s4b + kWave);
  __shared__ int lstk[kLdsStack];    // node stack; deeper entries go to the global region
  const int lane = threadIdx.x;
  int* stk = a.stacks + (size_t) blockIdx.x*kStackCap;
  const int N = a.N;
  const KernelDesc& K = a.p.kern;
  const double kr = K.range, krsqd = K.rangesqd;
  GroupQueue gq = group_queue(a.glist ? a.glist[a.ncap] : a.ngroups);
  for (;;) {
    const int k = group_next(gq, a.counter);
    if (k >= gq.nq) break;
    const int g = a.glist ? a.glist[k] : k;
    const int G = a.grp_node[g], gfirst = a.grp_first[g], gcnt = a.grp_count[g];
    const bool has = lane < gcnt && (!a.act || a.act[gfirst + lane]);
    const int i = gfirst + lane;
    const int nact = __popcll(__ballot(has));
    if (nact == 0) {                 // no active target in this group
      if (lane == 0) { a.gnc[g] = 0; a.gnn[g] = 0; }
      continue;
    }
    const double4 gc = a.geo_c[G], gh = a.geo_h[G];
    const double hmaxG = gc.w;
    // fast multipoles: expansion point (group centre of mass, relative to the group centre), the
    // squared distance beyond which a cell's series converges for the whole group, and this
    // lane's share of the coefficients
    double rcx = 0.0, rcy = 0.0, rcz = 0.0, rg2t = 0.0;
    double fc[10];
    if (FAST) {
#pragma clang fp contract(off)
      const double* mG = a.mp + 10*(size_t) G;
      rcx = mG[0] - gc.x; rcy = mG[1] - gc.y; rcz = mG[2] - gc.z;
      const double ex = fabs(rcx) + gh.x, ey = fabs(rcy) + gh.y, ez = fabs(rcz) + gh.z;
      rg2t = (ex*ex + ey*ey + ez*ez)/a.p.thetamaxsqd;
      for (int k = 0; k < 10; k++) fc[k] = 0.0;
    }
    const bool mono = !PLUMMER && a.p.multipole == GDF_MULTIPOLE_MONOPOLE;
    const bool gmac = !PLUMMER && a.p.gravity_mac != GDF_MAC_GEOMETRIC;
    const double gfac = gmac ? a.gmac[g] : 0.0;
    double4 me = make_double4(0, 0, 0, 0), mr = make_double4(1, 0, 0, 0);
    if (has) { me = a.posm[i]; mr = a.grec[i]; }
    const double r2i = range2_rn(krsqd, mr.x);
    const float r2if = (float) r2i;
    // accumulators; gpot starts at the self term (m/h) wpot(0) (GradhSphTree.cpp:512)
    double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
    KahanF fx, fy, fz, fp;
    const double pot0 = (has && !PLUMMER) ? (me.w/mr.x)*kwpot<double>(K, 0.0) : 0.0;
    const double eps2 = a.p.eps2;
    const float eps2f = (float) eps2;
    // target position relative to the group centre (cell-relative frame of the FP32 path)
    const float rxi = (float) (me.x - gc.x), ryi = (float) (me.y - gc.y), rzi = (float) (me.z - gc.z);
    int sp = 1, phead = 0, ptail = 0, chead = 0, ctail = 0;
    int nnear = 0, ncell = 0;
    bool overflow = false;
    if (lane == 0) lstk[0] = a.grp_root ? a.grp_root[g] : 0;   // the group's segment subtree
    __syncthreads();
    while (sp > 0) {
      const int b = min(sp, kWave);
      int node = -1;
      if (lane < b) node = stack_get(lstk, stk, sp - 1 - lane);
      sp -= b;
      bool open = false, near = false, cell = false;
      int ef = 0, ec = 0, cl = 0, cr = 0;
      if (node >= 0) {
#pragma clang fp contract(off)
        // box, range and children in one round of loads
        const double4 nc = a.geo_c[node], nh = a.geo_h[node];
        NodeLinks nl;
        load_links(node, N, a.lnk, nl);
        double dx = nc.x - gc.x, dy = nc.y - gc.y, dz = nc.z - gc.z;
        nearest_image<double>(a.p, dx, dy, dz);
        const double gx = fmax(fabs(dx) - nh.x - gh.x, 0.0);
        const double gy = fmax(fabs(dy) - nh.y - gh.y, 0.0);
        const double gz = fmax(fabs(dz) - nh.z - gh.z, 0.0);
        const double dbox2 = gx*gx + gy*gy + gz*gz;
        const double hr = kr*fmax(hmaxG, nc.w);
        ef = nl.first; ec = nl.count;
        const bool bucket = ec <= a.nleaf || node >= N - 1;
        if (dbox2 < hr*hr*kGravNeibSlack) {
          if (bucket) near = true; else open = true;
        } else {
          const double ex = fmax(fabs(dx) - gh.x, 0.0);
          const double ey = fmax(fabs(dy) - gh.y, 0.0);
          const double ez = fmax(fabs(dz) - gh.z, 0.0);
          const double dmac2 = ex*ex + ey*ey + ez*ez;
          bool accept = dmac2 >= nh.w;
          if (gmac && accept) accept = !mac_opens(a.p, nh, a.mp + 10*(size_t) node, dmac2, gfac);
          if (FAST && accept) {
            const double* mc = a.mp + 10*(size_t) node;
            double cx = mc[0] - gc.x, cy = mc[1] - gc.y, cz = mc[2] - gc.z;
            nearest_image<double>(a.p, cx, cy, cz);
            cx -= rcx; cy -= rcy; cz -= rcz;
            accept = cx*cx + cy*cy + cz*cz >= rg2t;
          }
          if (accept) {
            if (ec == 1) near = true; else cell = true;
          } else {
            if (bucket) near = true; else open = true;
          }
        }
        if (open) { cl = nl.left; cr = nl.right; }
      }
      // children onto the stack
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
      // accepted cells
      const unsigned long long cm = __ballot(cell);
      if (cell) cbuf[(ctail + lanes_below(cm)) & (CCAP - 1)] = node;
      ctail += __popcll(cm);
      ncell += __popcll(cm);
      // direct particles
      const int ecount = near ? ec : 0;
      const int incl = wave_incl_scan(ecount);
      const int tot = __shfl(incl, 63, 64);
      for (int k = 0; k < ecount; k++) pbuf[(ptail + incl - ecount + k) & (PCAP - 1)] = ef + k;
      ptail += tot;
      nnear += tot;
      __syncthreads();
      // evaluate full batches (and everything left once the walk is done)
      while (ctail - chead >= kWave || (sp == 0 && ctail > chead)) {
        const int nb = min(ctail - chead, kWave);
        if (lane < nb) {
          const int c = cbuf[(chead + lane) & (CCAP - 1)];
          const double* m = a.mp + 10*(size_t) c;
          double cx = m[0], cy = m[1], cz = m[2];
          if (FAST) {
#pragma clang fp contract(off)
            // the cell relative to the expansion point, with the node test's replica
            double dx = cx - gc.x, dy = cy - gc.y, dz = cz - gc.z;
            nearest_image<double>(a.p, dx, dy, dz);
            fast_add(fc, dx - rcx, dy - rcy, dz - rcz, m[3], m + 4,
                     a.p.multipole == GDF_MULTIPOLE_FAST_QUADRUPOLE);
          }
          if (a.p.periodic) {
            // nearest periodic replica of the cell relative to the group centre
            double dx = cx - gc.x, dy = cy - gc.y, dz = cz - gc.z;
            nearest_image<double>(a.p, dx, dy, dz);
            cx = gc.x + dx; cy = gc.y + dy; cz = gc.z + dz;
          }
          const double qs = mono ? 0.0 : 1.0;      // monopole: no quadrupole moments
          if (MIXED) {
            s4a[lane] = make_float4((float) (cx - gc.x), (float) (cy - gc.y),
                                    (float) (cz - gc.z), (float) m[3]);
            s4b[lane] = make_float4((float) (qs*m[4]), (float) (qs*m[5]), (float) (qs*m[6]),
                                    (float) (qs*m[7]));
            s1[lane] = (float) (qs*m[8]);
          } else {
            sd[0][lane] = cx; sd[1][lane] = cy; sd[2][lane] = cz; sd[3][lane] = m[3];
            for (int k = 4; k < 9; k++) sd[k][lane] = qs*m[k];
          }
        }
        __syncthreads();
        // fast multipoles leave only the periodic corrections to the per-target sweep
        if (has && (!FAST || EWALD)) {
          if (MIXED) {
            float bx = 0.f, by = 0.f, bz = 0.f, bp = 0.f;
            #pragma unroll 4
            for (int q = 0; q < nb; q++) {
              const float4 c = s4a[q], qa = s4b[q];
              const float qq[5] = {qa.x, qa.y, qa.z, qa.w, s1[q]};
              if (!FAST)
                quad_term<float>(rxi - c.x, ryi - c.y, rzi - c.z, c.w, qq, bx, by, bz, bp,
                                 PLUMMER ? eps2f : 1.0e-20f);
              if (EWALD) ewald_corr<float>(a.p, c.x - rxi, c.y - ryi, c.z - rzi, c.w, bx, by, bz,
                                           bp);
            }
            fx.add(bx); fy.add(by); fz.add(bz); fp.add(bp);
          } else {
            for (int q = 0; q < nb; q++) {
              const double qq[5] = {sd[4][q], sd[5][q], sd[6][q], sd[7][q], sd[8][q]};
              if (!FAST)
                quad_term<double>(me.x - sd[0][q], me.y - sd[1][q], me.z - sd[2][q], sd[3][q],
                                  qq, ax, ay, az, pot, PLUMMER ? eps2 : 1.0e-20);
              if (EWALD) ewald_corr<double>(a.p, sd[0][q] - me.x, sd[1][q] - me.y,
                                            sd[2][q] - me.z, sd[3][q], ax, ay, az, pot);
            }
          }
        }
        chead += nb;
        __syncthreads();
      }
      while (ptail - phead >= kWave || (sp == 0 && ptail > phead)) {
        const int nb = min(ptail - phead, kWave);
        if (lane < nb) {
          const int j = pbuf[(phead + lane) & (PCAP - 1)];
          const double4 pj = a.posm[j];
          const double4 rj = a.grec[j];
          sidx[lane] = j;
          if (MIXED) {
            double dx = pj.x - gc.x, dy = pj.y - gc.y, dz = pj.z - gc.z;
            nearest_image<double>(a.p, dx, dy, dz);
            s4a[lane] = make_float4((float) dx, (float) dy, (float) dz, (float) pj.w);
            s1[lane] = (float) range2_rn(krsqd, rj.x);
            if (SOFT) s4b[lane] = make_float4((float) rj.x, (float) rj.y, (float) rj.z, 0.f);
          } else {
            sd[0][lane] = pj.x; sd[1][lane] = pj.y; sd[2][lane] = pj.z; sd[3][lane] = pj.w;
            sd[4][lane] = range2_rn(krsqd, rj.x);
            if (SOFT) { sd[5][lane] = rj.x; sd[6][lane] = rj.y; sd[7][lane] = rj.z; }
          }
        }
        __syncthreads();
        if (has) {
          if (MIXED) {
            const float hi = (float) mr.x, zi = (float) mr.y, hfi = (float) mr.z;
            float bx = 0.f, by = 0.f, bz = 0.f, bp = 0.f;
            // unrolled so that several LDS reads are in flight (same summation order)
            #pragma unroll 8
            for (int q = 0; q < nb; q++) {
              // every record is loaded whatever the pair's branch: no LDS read under a mask
              const float4 s = s4a[q];
              const float r2j = s1[q];
              const float dx = s.x - rxi, dy = s.y - ryi, dz = s.z - rzi;
              const float d2 = dot3_f(dx, dy, dz);
              const bool soft = (d2 < r2if) | (d2 < r2j);
              if (PLUMMER) {
                const float inv = rsq_f(d2 + eps2f);
                const float mj = sidx[q] == i ? 0.f : s.w;
                const float minv3 = mj*inv*inv*inv;
                bx += minv3*dx; by += minv3*dy; bz += minv3*dz; bp += mj*inv;
              } else if (SOFT) {
                const float4 hq = s4b[q];
                pair_fp32(K, d2, soft, dx, dy, dz, s.w, hi, hq.x, zi, hq.y, hfi, hq.z, bx, by,
                          bz, bp);
              } else {
                const float inv = rsq_f(d2 + 1.0e-20f);
                const float mj = soft ? 0.f : s.w;
                const float minv3 = mj*inv*inv*inv;
                bx += minv3*dx; by += minv3*dy; bz += minv3*dz; bp += mj*inv;
              }
              if (EWALD) ewald_corr<float>(a.p, dx, dy, dz, s.w, bx, by, bz, bp);
            }
            fx.add(bx); fy.add(by); fz.add(bz); fp.add(bp);
          } else {
            for (int q = 0; q < nb; q++) {
              double dx = sd[0][q] - me.x, dy = sd[1][q] - me.y, dz = sd[2][q] - me.z;
              nearest_image<double>(a.p, dx, dy, dz);
              if (PLUMMER) {
                const double inv = rsq_d(dot3_rn(dx, dy, dz) + eps2);
                const double mj = sidx[q] == i ? 0.0 : sd[3][q];
                const double minv3 = mj*inv*inv*inv;
                ax += minv3*dx; ay += minv3*dy; az += minv3*dz; pot += mj*inv;
              } else if (SOFT) {
                pair_fp64(K, krsqd, dx, dy, dz, sd[3][q], mr.x, sd[5][q], mr.y, sd[6][q], mr.z,
                          sd[7][q], ax, ay, az, pot);
              } else {
                const double d2 = dot3_rn(dx, dy, dz);
                const bool soft = (d2 < r2i) | (d2 < sd[4][q]);
                const double inv = rsq_d(d2 + 1.0e-20);
                const double mj = soft ? 0.0 : sd[3][q];
                const double minv3 = mj*inv*inv*inv;
                ax += minv3*dx; ay += minv3*dy; az += minv3*dz; pot += mj*inv;
              }
              if (EWALD) ewald_corr<double>(a.p, dx, dy, dz, sd[3][q], ax, ay, az, pot);
            }
          }
        }
        phead += nb;
        __syncthreads();
      }
    }
    if (overflow) {
      if (lane == 0) atomicOr(&a.flags[2], 1);
      continue;
    }
    if (FAST) {
      // the lanes' coefficient sums, reduced by a butterfly (every lane ends with the same sums)
      for (int k = 0; k < 10; k++) {
        double v = fc[k];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        fc[k] = v;
      }
    }
    if (has) {
      if (MIXED) { ax = (double) fx.s; ay = (double) fy.s; az = (double) fz.s; pot = (double) fp.s; }
      if (FAST) {
        // ApplyForcesTaylor (NeighbourSearch.h:735-760), first order about the expansion point
        const double* mG = a.mp + 10*(size_t) G;
        const double d0 = me.x - mG[0], d1 = me.y - mG[1], d2 = me.z - mG[2];
        ax += fc[1] + fc[4]*d0 + fc[5]*d1 + fc[7]*d2;
        ay += fc[2] + fc[5]*d0 + fc[6]*d1 + fc[8]*d2;
        az += fc[3] + fc[7]*d0 + fc[8]*d1 + fc[9]*d2;
        pot += fc[0] + fc[1]*d0 + fc[2]*d1 + fc[3]*d2;
      }
      if (SOFT) {
        a.atx[i] = ax; a.aty[i] = ay; a.atz[i] = az;
        a.gpot[i] = pot0 + pot;
      } else if (a.gwx) {
        a.gwx[i] = ax; a.gwy[i] = ay; a.gwz[i] = az;
        a.gwp[i] = pot; a.gwp0[i] = pot0;
      } else {
        a.atx[i] += ax; a.aty[i] += ay; a.atz[i] += az;
        a.gpot[i] = pot0 + a.gpot[i] + pot;
      }
    }
    if (lane == 0) {
      a.gnc[g] = ncell;
      a.gnn[g] = nnear;
      atomicAdd((unsigned long long*) &a.icount[2], (unsigned long long) nnear*nact);
      atomicAdd((unsigned long long*) &a.icount[3], (unsigned long long) ncell*(FAST ? 1 : nact));
    }
  }
}

struct NbDirectArgs {
  int N;
  const double* x; const double* y; const double* z;
  const double* vx; const double* vy; const double* vz;
  const double* m; const double* h;
  double* ax; double* ay; double* az; double* adx; double* ady; double* adz; double* gpot;
  KernelDesc kern;
  int add;          // add to a, adot and gpot (after the stars' gas terms) instead of writing
};

// Star-star direct sums, one thread per star, source stars staged 256 at a time in LDS.
// SMOOTH = false: Nbody::CalculateDirectGravForces (src/Nbody/Nbody.cpp:233-293), point
// masses with jerk; SMOOTH = true: NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces
// (src/Nbody/NbodyLeapfrogKDK.cpp:78-150), kernel softening over the mean smoothing length,
// whose (2/(h_i+h_j))^3 factor the reference forms in single precision (powf).
template <bool SMOOTH>
__global__ __launch_bounds__(256) void k_nbody_direct(NbDirectArgs a) {
  constexpr int T = 256;
  __shared__ double sx[T], sy[T], sz[T], svx[T], svy[T], svz[T], sm[T], sh[T];
  const int i = blockIdx.x*T + threadIdx.x;
  const bool has = i < a.N;
  double xi = 0, yi = 0, zi = 0, vxi = 0, vyi = 0, vzi = 0, hi = 1.0;
  if (has) {
    xi = a.x[i]; yi = a.y[i]; zi = a.z[i]; vxi = a.vx[i]; vyi = a.vy[i]; vzi = a.vz[i];
    if (SMOOTH) hi = a.h[i];
  }
  double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0, pot = 0;
  for (int j0 = 0; j0 < a.N; j0 += T) {
    const int jl = j0 + threadIdx.x;
    if (jl < a.N) {
      sx[threadIdx.x] = a.x[jl]; sy[threadIdx.x] = a.y[jl]; sz[threadIdx.x] = a.z[jl];
      svx[threadIdx.x] = a.vx[jl]; svy[threadIdx.x] = a.vy[jl]; svz[threadIdx.x] = a.vz[jl];
      sm[threadIdx.x] = a.m[jl]; sh[threadIdx.x] = SMOOTH ? a.h[jl] : 0.0;
    }
    __syncthreads();
    const int nb = min(T, a.N - j0);
    if (has) {
      for (int q = 0; q < nb; q++) {
        if (j0 + q == i) continue;
        const double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
        const double dvx = svx[q] - vxi, dvy = svy[q] - vyi, dvz = svz[q] - vzi;
        const double drsqd = dx*dx + dy*dy + dz*dz;
        const double mj = sm[q];
        if (!SMOOTH) {
          const double invdrmag = 1.0/sqrt(drsqd);
          const double drdt = (dvx*dx + dvy*dy + dvz*dz)*invdrmag;
          const double inv3 = invdrmag*invdrmag*invdrmag;
          pot += mj*invdrmag;
          ax += mj*dx*inv3; ay += mj*dy*inv3; az += mj*dz*inv3;
          const double c3 = 3.0*drdt*invdrmag;
          jx += mj*inv3*(dvx - c3*dx); jy += mj*inv3*(dvy - c3*dy); jz += mj*inv3*(dvz - c3*dz);
        } else {
          const double drmag = sqrt(drsqd) + 1.0e-20;
          const double invdrmag = 1.0/drmag;
          const double invhmean = 2.0/(hi + sh[q]);
          const double drdt = (dvx*dx + dvy*dy + dvz*dz)*invdrmag;
          const double s = drmag*invhmean;
          const double paux = mj*invhmean*invhmean*kwgrav<double>(a.kern, s)*invdrmag;
          // the reference's single-precision powf(invhmean, 3): the correctly rounded cube
          // (exact in double before the one rounding to float), which is what glibc's powf
          // returns for all but ~0.07% of arguments (those differ by one float ulp)
          const float hmf = (float) invhmean;
          const double wmean = kw0<double>(a.kern, s)*
                               (double) (float) ((double) hmf*(double) hmf*(double) hmf);
          pot += mj*invhmean*kwpot<double>(a.kern, s);
          ax += paux*dx; ay += paux*dy; az += paux*dz;
          const double c3 = 3.0*paux*drdt*invdrmag;
          const double cw = 2.0*6.28318530717959*mj*drdt*wmean*invdrmag;
          jx += paux*dvx - c3*dx + cw*dx;
          jy += paux*dvy - c3*dy + cw*dy;
          jz += paux*dvz - c3*dz + cw*dz;
        }
      }
    }
    __syncthreads();
  }
  if (!has) return;
  if (a.add) {
    a.ax[i] += ax; a.ay[i] += ay; a.az[i] += az;
    a.adx[i] += jx; a.ady[i] += jy; a.adz[i] += jz;
    a.gpot[i] += pot;
    return;
  }
  a.ax[i] = ax; a.ay[i] = ay; a.az[i] = az;
  a.adx[i] = jx; a.ady[i] = jy; a.adz[i] = jz;
  a.gpot[i] = pot;
}

__global__ void k_pack_grav(int N, const double* __restrict__ h, const double* __restrict__ zeta,
                            const double* __restrict__ hf, double4* __restrict__ g) {
  const int i = blockIdx.x*256 + threadIdx.x;
  if (i < N) g[i] = make_double4(h[i], zeta[i], hf[i], 0.0);
}

__global__ void k_zero8(int* p) { if (threadIdx.x < 8) p[threadIdx.x] = 0; }

// the walk's sums joined to the hydro pass's in the order of the single-stream update
// (atx += walk; gpot = self + gpot + walk), then a += atree (GradhSphTree.cpp:615-617)
struct JoinArgs {
  int N;
  const int* act;
  const double* gw[3]; const double* gwp; const double* gwp0;
  double* at[3]; double* gpot; double* a[3];
};
__global__ void k_grav_join(JoinArgs j) {
  const int i = blockIdx.x*256 + threadIdx.x;
  if (i >= j.N || (j.act && !j.act[i])) return;
  for (int k = 0; k < 3; k++) {
    const double t = j.at[k][i] + j.gw[k][i];
    j.at[k][i] = t;
    j.a[k][i] += t;
  }
  j.gpot[i] = j.gwp0[i] + j.gpot[i] + j.gwp[i];
}

}  // namespace

int nbody_direct(gdf_ctx* c, int smoothed) {
  const int N = c->N;
  if (N <= 0) return fail(c, GDF_ERR_INVALID, "no particles");
  NbDirectArgs a;
  a.N = N;
  a.x = c->pf[F_X]; a.y = c->pf[F_Y]; a.z = c->pf[F_Z];
  a.vx = c->pf[F_VX]; a.vy = c->pf[F_VY]; a.vz = c->pf[F_VZ];
  a.m = c->pf[F_M]; a.h = c->pf[F_H];
  a.ax = c->df[D_AX]; a.ay = c->df[D_AY]; a.az = c->df[D_AZ];
  a.adx = c->df[D_ADOTX]; a.ady = c->df[D_ADOTY]; a.adz = c->df[D_ADOTZ];
  a.gpot = c->df[D_GPOT];
  a.kern = c->phys.kern;
  a.add = 0;
  const int nb = (N + 255)/256;
  if (smoothed) k_nbody_direct<true><<<nb, 256, 0, c->stream>>>(a);
  else k_nbody_direct<false><<<nb, 256, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int nbody_direct_stars(gdf_ctx* c, int smoothed) {
  const int N = c->Nstar;
  if (N <= 0) return GDF_OK;
  NbDirectArgs a;
  a.N = N;
  a.x = c->sv[S_X]; a.y = c->sv[S_Y]; a.z = c->sv[S_Z];
  a.vx = c->sv[S_VX]; a.vy = c->sv[S_VY]; a.vz = c->sv[S_VZ];
  a.m = c->sv[S_M]; a.h = c->sv[S_H];
  a.ax = c->sv[S_AX]; a.ay = c->sv[S_AY]; a.az = c->sv[S_AZ];
  a.adx = c->sv[S_ADX]; a.ady = c->sv[S_ADY]; a.adz = c->sv[S_ADZ];
  a.gpot = c->sv[S_GPOT];
  a.kern = c->phys.kern;
  a.add = 1;
  const int nb = (N + 255)/256;
  if (smoothed) k_nbody_direct<true><<<nb, 256, 0, c->stream>>>(a);
  else k_nbody_direct<false><<<nb, 256, 0, c->stream>>>(a);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

namespace {

// gas walks: the fast-multipole variant when the multipole option asks for it
template <bool MIXED, bool SOFT, bool EWALD>
void launch_gas(const GravArgs& a, int grid, hipStream_t s) {
  if (a.p.multipole >= GDF_MULTIPOLE_FAST_MONOPOLE)
    k_gravity<MIXED, SOFT, false, EWALD, true><<<grid, kWave, 0, s>>>(a);
  else
    k_gravity<MIXED, SOFT, false, EWALD, false><<<grid, kWave, 0, s>>>(a);
}

GravArgs grav_args(gdf_ctx* c) {
  GravArgs a;
  a.N = c->N; a.ngroups = c->ngroups; a.nleaf = c->phys.nleaf;
  a.grp_node = c->grp_node; a.grp_first = c->grp_first; a.grp_count = c->grp_count;
  a.lnk = c->lnk;
  a.geo_c = c->geo_c; a.geo_h = c->geo_h; a.mp = c->mp;
  a.posm = c->posm; a.grec = c->grec;
  a.atx = c->df[D_ATX]; a.aty = c->df[D_ATY]; a.atz = c->df[D_ATZ]; a.gpot = c->df[D_GPOT];
  a.stacks = c->stacks; a.counter = c->counters + 4; a.flags = c->counters;
  a.icount = c->icount; a.gnc = c->grav_nc; a.gnn = c->grav_nn;
  a.act = c->nlevels > 1 ? c->pi[I_ACT] : nullptr;
  a.grp_root = c->nseg > 0 ? c->grp_root : nullptr;
  a.gmac = c->grp_gmac;
  a.gwx = a.gwy = a.gwz = a.gwp = a.gwp0 = nullptr;
  a.glist = nullptr; a.ncap = c->Ncap;
  a.p = c->phys;
  return a;
}

// in a block step, the compacted list of groups with active particles (built on the main stream)
// zero[0, nzero): the flags and queue counters of the walk, cleared with the list (or by the
// caller when there is no list: returns false)
bool grav_group_list(gdf_ctx* c, GravArgs& a, int* zero, int nzero, int& rc) {
  rc = GDF_OK;
  if (!a.act || !c->glist) return false;
  int* gl = c->glist + 2*((size_t) c->Ncap + 1);
  rc = group_list(c, a.act, gl, nullptr, c->grav_nc, c->grav_nn, zero, nzero);
  a.glist = gl;
  return true;
}

}  // namespace

int gravity_launch(gdf_ctx* c) {
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "gravity pass before tree build");
  if (!c->stacks2 || !c->phys.hydro_forces || c->phys.nbody)
    return fail(c, GDF_ERR_INVALID, "gravity walk beside the hydro pass needs hydro + self-gravity");
  const int N = c->N;
  // on the second stream beside the scatter-gather walk and the hydro pass: measured faster
  // than running them in turn, both at 1M active particles and in sparse block steps
  hipStream_t s2 = c->stream2;
  GravArgs a = grav_args(c);
  int rc = GDF_OK;
  const bool listed = grav_group_list(c, a, c->counters + 16, 16, rc);   // flags and queue
  GDF_TRY(rc);
  GDF_HIP(c, hipEventRecord(c->ev_dens, c->stream));        // h, zeta, tree, posm are ready
  GDF_HIP(c, hipStreamWaitEvent(s2, c->ev_dens, 0));
  k_pack_grav<<<nblocks(N, 256), 256, 0, s2>>>(N, c->pf[F_H], c->df[D_ZETA], c->df[D_HFACTOR],
                                                 c->grec);
  a.stacks = c->stacks2; a.counter = c->counters + 24; a.flags = c->counters + 16;
  a.gwx = c->df[D_GWX]; a.gwy = c->df[D_GWY]; a.gwz = c->df[D_GWZ];
  a.gwp = c->df[D_GWP]; a.gwp0 = c->df[D_GWP0];
  if (!listed) {
    k_zero8<<<1, 64, 0, s2>>>(c->counters + 24);
    k_zero8<<<1, 64, 0, s2>>>(c->counters + 16);
  }
  // half the persistent wavefronts of a lone walk: at the full count they fill every CU's LDS
  // and the scatter-gather walk beside them waits for the gravity walk to drain
  const int grid = std::min(c->ngroups, kGravSlots/2);
  const int kt = ktime_begin(c, s2, 5);
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  if (c->phys.ewald) {
    if (mixed) launch_gas<true, false, true>(a, grid, s2);
    else launch_gas<false, false, true>(a, grid, s2);
  } else {
    if (mixed) launch_gas<true, false, false>(a, grid, s2);
    else launch_gas<false, false, false>(a, grid, s2);
  }
  GDF_HIP(c, hipGetLastError());
  ktime_end(c, s2, kt);
  GDF_HIP(c, hipMemcpyAsync(c->h_counters + 16, c->counters + 16, sizeof(int)*8,
                            hipMemcpyDeviceToHost, s2));
  GDF_HIP(c, hipEventRecord(c->ev_grav, s2));
  return GDF_OK;
}

int gravity_join(gdf_ctx* c) {
  GDF_HIP(c, hipStreamWaitEvent(c->stream, c->ev_grav, 0));
  JoinArgs j;
  j.N = c->N; j.act = c->nlevels > 1 ? c->pi[I_ACT] : nullptr;
  for (int k = 0; k < 3; k++) {
    j.gw[k] = c->df[D_GWX + k]; j.at[k] = c->df[D_ATX + k]; j.a[k] = c->df[D_AX + k];
  }
  j.gwp = c->df[D_GWP]; j.gwp0 = c->df[D_GWP0]; j.gpot = c->df[D_GPOT];
  k_grav_join<<<nblocks(c->N, 256), 256, 0, c->stream>>>(j);
  GDF_HIP(c, hipGetLastError());
  GDF_HIP(c, hipEventSynchronize(c->ev_grav));
  if (c->h_counters[16 + 2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in gravity pass");
  return GDF_OK;
}

int gravity_walk(gdf_ctx* c) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "gravity pass before tree build");
  k_pack_grav<<<nblocks(N, 256), 256, 0, s>>>(N, c->pf[F_H], c->df[D_ZETA], c->df[D_HFACTOR],
                                                c->grec);
  GravArgs a = grav_args(c);
  int rc = GDF_OK;
  if (!grav_group_list(c, a, c->counters, 12, rc)) {
    k_zero8<<<1, 64, 0, s>>>(c->counters + 4);
    k_zero8<<<1, 64, 0, s>>>(c->counters);
  }
  GDF_TRY(rc);
  const int grid = std::min(c->ngroups, kGravSlots);
  const int kt = ktime_begin(c, s, 5);
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  if (c->phys.nbody) {
    if (mixed) k_gravity<true, true, true><<<grid, kWave, 0, s>>>(a);
    else k_gravity<false, true, true><<<grid, kWave, 0, s>>>(a);
  } else if (c->phys.hydro_forces) {
    if (c->phys.ewald) {
      if (mixed) launch_gas<true, false, true>(a, grid, s);
      else launch_gas<false, false, true>(a, grid, s);
    } else {
      if (mixed) launch_gas<true, false, false>(a, grid, s);
      else launch_gas<false, false, false>(a, grid, s);
    }
  } else if (c->phys.ewald) {
    if (mixed) launch_gas<true, true, true>(a, grid, s);
    else launch_gas<false, true, true>(a, grid, s);
  } else {
    if (mixed) launch_gas<true, true, false>(a, grid, s);
    else launch_gas<false, true, false>(a, grid, s);
  }
  GDF_HIP(c, hipGetLastError());
  ktime_end(c, s, kt);
  GDF_HIP(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(int)*8, hipMemcpyDeviceToHost, s));
  GDF_HIP(c, hipStreamSynchronize(s));
  if (c->h_counters[2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in gravity pass");
  return GDF_OK;
}

}  // namespace gdf
