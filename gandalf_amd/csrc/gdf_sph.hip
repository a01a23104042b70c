// grad-h SPH passes: neighbour-list walks, h-rho iteration and hydro forces.
//
// Replaces the OpenMP cell loops of GradhSphTree::UpdateAllSphProperties
// (src/GradhSph/GradhSphTree.cpp:83-271) and UpdateAllSphHydroForces (:280-435) with the
// per-particle physics of GradhSph::ComputeH (src/GradhSph/GradhSph.cpp:142-326) and
// ComputeSphHydroForces / ComputeSphHydroGravForces (:361-460, :474-585).
//
// Work mapping: one wavefront per walk group (a tree node of <= 64 Morton-consecutive
// particles).  The wave walks the tree once for all its targets, stages candidate particle
// positions in LDS in batches of 64 and every lane tests its own target against the batch,
// appending exact neighbours to an interleaved list (entry k of particle i at k*N + i).  The
// pair passes then run one thread per particle over its list with full SIMD occupancy and
// reuse the list across h-rho iterations.
#include <hip/hip_runtime.h>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;
constexpr int CANDCAP = 1024;              // ring buffer of candidate slots (power of two)
constexpr double NODE_SLACK = 1.0 + 1e-12; // node tests err on the inclusive side

struct ListWalk {
  int N, ngroups, cap, nleaf, mode;         // mode 0 search (k*hs_i, inclusive), 1 scatter-gather
  int mixed;                                // scatter-gather of the mixed path: FP32 group frame
  const int* grp_node; const int* grp_first; const int* grp_count;
  const int* first; const int* last; const int* left; const int* right;
  const double4* geo_c; const double4* geo_h;
  const double* x; const double* y; const double* z; const double* hcand; const double* htarget;
  int* list; int* nn; int* stacks; int* counter; int* flags;
  PhysParams p;
};

__device__ __forceinline__ void node_range(int n, int N, const int* first, const int* last,
                                           int& f, int& cnt) {
  if (n >= N - 1) { f = n - (N - 1); cnt = 1; }
  else { f = first[n]; cnt = last[n] - first[n] + 1; }
}

// MODE 0: search set |r_ij|^2 <= (k hs_i)^2 (FP64); 1: scatter-gather set
// |r_ij|^2 < (k h_i)^2 or < (k h_j)^2 (FP64, the oracle's operation order); 2: the same set
// decided in FP32 in the group-centred frame, exactly as the mixed-precision gravity walk
// decides it.  PER: periodic nearest images.
template <int MODE, bool PER>
__global__ __launch_bounds__(64) void k_build_lists(ListWalk a) {
  __shared__ int cand[CANDCAP];
  __shared__ double4 sc[kWave];      // MODE 0/1: (x, y, z, (k h_j)^2)
  __shared__ float4 fc[kWave];       // MODE 2: group-frame (x, y, z, (k h_j)^2)
  __shared__ int sj[kWave];
  const int lane = threadIdx.x;
  int* stk = a.stacks + (size_t) blockIdx.x*kStackCap;
  const int N = a.N;
  const double krsqd = a.p.kern.rangesqd, kr = a.p.kern.range;
  PhysParams pp = a.p;
  pp.periodic = PER ? 1 : 0;
  int range = blockIdx.x % kXcds;
  for (;;) {
    const int g = wave_next_group(a.counter, a.ngroups, range);
    if (g >= a.ngroups) break;
    const int G = a.grp_node[g], gfirst = a.grp_first[g], gcnt = a.grp_count[g];
    const bool has = lane < gcnt;
    const int i = gfirst + lane;
    double xi = 0.0, yi = 0.0, zi = 0.0, hi = 0.0, r2i = 0.0;
    if (has) {
      xi = a.x[i]; yi = a.y[i]; zi = a.z[i]; hi = a.htarget[i];
      r2i = range2_rn(krsqd, hi);
    }
    const double hG = wave_max_d(has ? hi : 0.0);
    const double4 gc = a.geo_c[G], gh = a.geo_h[G];
    const float rxi = (float) (xi - gc.x), ryi = (float) (yi - gc.y), rzi = (float) (zi - gc.z);
    const float r2if = (float) r2i;
    int count = 0;
    int sp = 1, head = 0, tail = 0;
    bool overflow = false;
    if (lane == 0) stk[0] = 0;
    __threadfence_block();
    while (sp > 0) {
      const int b = min(sp, kWave);
      int node = -1;
      if (lane < b)
        node = stk[sp - 1 - lane];
      sp -= b;
      bool open = false, emit = false;
      int ef = 0, ec = 0, cl = 0, cr = 0;
      if (node >= 0) {
        const double4 nc = a.geo_c[node], nh = a.geo_h[node];
        double dx = nc.x - gc.x, dy = nc.y - gc.y, dz = nc.z - gc.z;
        nearest_image<double>(pp, dx, dy, dz);
        const double gx = fmax(fabs(dx) - nh.x - gh.x, 0.0);
        const double gy = fmax(fabs(dy) - nh.y - gh.y, 0.0);
        const double gz = fmax(fabs(dz) - nh.z - gh.z, 0.0);
        const double dbox2 = gx*gx + gy*gy + gz*gz;
        const double rr = MODE == 0 ? kr*hG : kr*fmax(hG, nc.w);
        if (dbox2 <= rr*rr*NODE_SLACK) {
          node_range(node, N, a.first, a.last, ef, ec);
          if (ec <= a.nleaf) emit = true;
          else { open = true; cl = a.left[node]; cr = a.right[node]; }
        }
      }
      const unsigned long long om = __ballot(open);
      const int nopen = __popcll(om);
      if (sp + 2*nopen > kStackCap) { overflow = true; break; }
      if (open) {
        const int o = sp + 2*lanes_below(om);
        stk[o] = cr;
        stk[o + 1] = cl;
      }
      sp += 2*nopen;
      const int ecount = emit ? ec : 0;
      const int incl = wave_incl_scan(ecount);
      const int tot = __shfl(incl, 63, 64);
      for (int k = 0; k < ecount; k++) cand[(tail + incl - ecount + k) & (CANDCAP - 1)] = ef + k;
      tail += tot;
      __syncthreads();
      // flush complete batches
      while (tail - head >= kWave || (sp == 0 && tail > head)) {
        const int nb = min(tail - head, kWave);
        if (lane < nb) {
          const int j = cand[(head + lane) & (CANDCAP - 1)];
          sj[lane] = j;
          const double r2j = range2_rn(krsqd, a.hcand[j]);
          if (MODE == 2) {
            double dx = a.x[j] - gc.x, dy = a.y[j] - gc.y, dz = a.z[j] - gc.z;
            nearest_image<double>(pp, dx, dy, dz);
            fc[lane] = make_float4((float) dx, (float) dy, (float) dz, (float) r2j);
          } else {
            sc[lane] = make_double4(a.x[j], a.y[j], a.z[j], r2j);
          }
        }
        __syncthreads();
        if (has) {
          for (int q = 0; q < nb; q++) {
            bool in;
            if (MODE == 2) {
              const float4 c = fc[q];
              const float d2 = dot3_f(c.x - rxi, c.y - ryi, c.z - rzi);
              in = (d2 < r2if) || (d2 < c.w);
            } else {
              const double4 c = sc[q];
              double dx = c.x - xi, dy = c.y - yi, dz = c.z - zi;
              nearest_image<double>(pp, dx, dy, dz);
              const double d2 = dot3_rn(dx, dy, dz);
              if (MODE == 0) in = d2 <= r2i;
              else in = (d2 < r2i) || (d2 < c.w);
            }
            if (in) {
              if (count < a.cap) a.list[(size_t) count*N + i] = sj[q];
              count++;
            }
          }
        }
        head += nb;
        __syncthreads();
      }
      __threadfence_block();
    }
    if (overflow) {
      if (lane == 0) atomicOr(&a.flags[2], 1);
      continue;
    }
    if (has) {
      a.nn[i] = count;
      if (count > a.cap) atomicMax(&a.flags[1], count);
    }
  }
}

struct DensArgs {
  int N, cap, nopbc;
  const int* list; const int* nn;
  const double4* posm;
  double* h; double* hsearch; double* rho; double* invomega; double* zeta; double* hfactor;
  double* u; double* sound; double* pressure;
  int* status; int* flags; long long* icount;
  PhysParams p;
};

__device__ __forceinline__ void wave_add_ll(long long* dst, long long v) {
  #pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (__lane_id() == 0 && v != 0) atomicAdd((unsigned long long*) dst, (unsigned long long) v);
}

// grad-h iteration of one particle, GradhSph.cpp:142-326 (fixed point, bisection after 30).
// Returns 1 done, 0 retry with a larger search radius, -1 no convergence.
template <typename T>
__device__ int density_one(const DensArgs& a, const PhysParams& p, int i, int n,
                           long long& ngather, long long& evals) {
  const double4 me = a.posm[i];
  const double m = me.w;
  const double hmax = a.hsearch[i];
  const double h_fac = p.h_fac, h_conv = p.h_converge;
  const double rsq = p.kern.rangesqd;
  double h = a.h[i];
  double h_lower = 0.0, h_upper = hmax;
  double rho = 0.0, invomega = 0.0, zeta = 0.0;
  int iteration = 0;
  do {
    iteration++;
    const double invh = 1.0/h;
    const double hfac3 = invh*invh*invh;
    const double invhsqd = invh*invh;
    int inside = 0;
    if (sizeof(T) == 8) {
      double r_ = 0.0, om_ = 0.0, z_ = 0.0;
      for (int k = 0; k < n; k++) {
        const int j = a.list[(size_t) k*a.N + i];
        const double4 pj = a.posm[j];
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
      rho = r_; invomega = om_; zeta = z_;
    } else {
      KahanF r_, om_, z_;
      const float invhf = (float) invh, invhsqdf = (float) invhsqd;
      for (int k = 0; k < n; k++) {
        const int j = a.list[(size_t) k*a.N + i];
        const double4 pj = a.posm[j];
        double ddx = pj.x - me.x, ddy = pj.y - me.y, ddz = pj.z - me.z;
        nearest_image<double>(p, ddx, ddy, ddz);
        const float dx = (float) ddx, dy = (float) ddy, dz = (float) ddz;
        const float ssqd = invhsqdf*(dx*dx + dy*dy + dz*dz);
        inside += ssqd < (float) rsq;
        float w0, wom, wz;
        kdens<float>(p.kern, ssqd, w0, wom, wz);
        const float mj = (float) pj.w;
        r_.add(mj*w0);
        om_.add(mj*invhf*wom);
        z_.add(mj*wz);
      }
      rho = (double) r_.s; invomega = (double) om_.s; zeta = (double) z_.s;
    }
    evals += n;
    ngather = inside;
    rho *= hfac3;
    invomega *= hfac3;
    zeta *= invhsqd;
    const double hrho = rho > 0.0 ? h_fac*pow(m/rho, 1.0/3) : 0.0;
    if (rho > 0.0 && h > h_lower && fabs(h - hrho)*invh < h_conv) break;
    if (iteration < 30) h = hrho;
    else if (iteration == 30) h = 0.5*(h_lower + h_upper);
    else if (iteration < 150) {
      if (rho < 1.0e-20 || h > hrho) h_upper = h; else h_lower = h;
      h = 0.5*(h_lower + h_upper);
    } else return -1;
    if (h > hmax) return 0;
  } while (h > h_lower && h < h_upper);
  // GradhSph.cpp:260-317: final h from rho, grad-h terms, EOS
  const double hrho = h_fac*pow(m/rho, 1.0/3);
  h = fmax(hrho, h_lower);
  const double invh = 1/h;
  const double hfac4 = invh*invh*invh*invh;
  const double deriv = -(1.0/3)*h/rho;
  double iom = 1.0 - deriv*invomega;
  iom = 1.0/iom;
  const double zt = deriv*zeta*iom;
  a.h[i] = h; a.rho[i] = rho; a.invomega[i] = iom; a.zeta[i] = zt; a.hfactor[i] = hfac4;
  double u = a.u[i], snd;
  if (p.eos == 1) { u = p.temp0/p.gammam1/p.mu_bar; snd = sqrt(p.gammam1*u); }
  else if (p.eos == 2) {
    u = p.temp0*(1.0 + pow(rho*p.invrho_bary, p.gammam1))/p.gammam1/p.mu_bar;
    snd = sqrt(p.gammam1*u);
  } else snd = sqrt(p.gamma*p.gammam1*u);
  a.u[i] = u; a.sound[i] = snd; a.pressure[i] = p.gammam1*rho*u;
  return 1;
}

template <typename T>
__global__ __launch_bounds__(BS) void k_density(DensArgs a) {
  const int i = xcd_block(blockIdx.x, gridDim.x)*BS + threadIdx.x;
  long long ng = 0, ev = 0;
  if (i < a.N && a.status[i] != 0) {
    const int n = a.nn[i];
    if (n > a.cap) {
      atomicMax(&a.flags[1], n);
    } else {
      PhysParams p = a.p;
      if (a.nopbc) p.periodic = 0;
      long long ngi = 0;
      const int rc = density_one<T>(a, p, i, n, ngi, ev);
      if (rc == 1) { a.status[i] = 0; ng = ngi; }
      else if (rc == 0) { a.hsearch[i] = a.hsearch[i]*1.1; atomicAdd(&a.flags[0], 1); }
      else { a.status[i] = 0; atomicOr(&a.flags[3], 1); }
    }
  }
  wave_add_ll(&a.icount[0], ng);
  wave_add_ll(&a.icount[4], ev);
}

__global__ void k_init_density(int N, const double* __restrict__ h, double* __restrict__ hs,
                               int* __restrict__ status) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  hs[i] = 1.1*h[i];
  status[i] = 1;
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
  int N, cap;
  const int* list; const int* nn;
  const double4* posm; const double4* A; const double4* B;
  const double* invomega; const double* pressure; const double* zeta;
  double* ax; double* ay; double* az; double* dudt; double* div_v;
  double* gx; double* gy; double* gz; double* gpot;    // softened neighbour gravity (VAR 1)
  int* flags; long long* icount;
  PhysParams p;
};

// VAR 0: ComputeSphHydroForces (GradhSph.cpp:361-460); VAR 1: ComputeSphHydroGravForces
// (:474-585), the hydro terms plus the kernel-softened gravity of the pair (into gx, gy, gz,
// gpot; the tree walk adds the direct and multipole terms afterwards)
template <typename T, int VAR>
__global__ __launch_bounds__(BS) void k_hydro(HydroArgs a) {
  const int i0 = xcd_block(blockIdx.x, gridDim.x)*BS + threadIdx.x;
  wave_add_ll(&a.icount[1], i0 < a.N ? (long long) a.nn[i0] : 0LL);
  const int i = i0;
  if (i >= a.N) return;
  const int n = a.nn[i];
  if (n > a.cap) { atomicMax(&a.flags[1], n); return; }
  const PhysParams& p = a.p;
  const double4 me = a.posm[i], mA = a.A[i], mB = a.B[i];
  const T invh_i = T(mA.w);
  const T invrho_i = T(mB.x);
  const T pterm_i = T(mB.y), snd_i = T(mB.z), hf_i = T(mB.w);
  const T alpha = T(p.alpha_visc), beta = T(p.beta_visc);
  const T invhsqd_i = invh_i*invh_i;
  const T zeta_i = VAR == 1 ? T(a.zeta[i]) : T(0);
  T ax = 0, ay = 0, az = 0, du = 0, dv = 0;
  T gx = 0, gy = 0, gz = 0, gp = 0;
  for (int k = 0; k < n; k++) {
    const int j = a.list[(size_t) k*a.N + i];
    const double4 pj = a.posm[j], Aj = a.A[j], Bj = a.B[j];
    double ddx = pj.x - me.x, ddy = pj.y - me.y, ddz = pj.z - me.z;
    nearest_image<double>(p, ddx, ddy, ddz);
    T dx = T(ddx), dy = T(ddy), dz = T(ddz);
    const T mj = T(pj.w);
    const T invh_j = T(Aj.w);
    const T invrho_j = T(Bj.x);
    T drmag, dvdr;
    if (VAR == 0) {
      drmag = sqrt(dx*dx + dy*dy + dz*dz);
      if (drmag > T(0)) { dx /= drmag; dy /= drmag; dz /= drmag; }
    } else {
      drmag = sqrt(dx*dx + dy*dy + dz*dz + T(1.0e-20));
      const T inv = T(1)/drmag;
      dx *= inv; dy *= inv; dz *= inv;
    }
    const T si = drmag*invh_i, sj = drmag*invh_j;
    const T wkerni = hf_i*kw1<T>(p.kern, si);
    const T wkernj = T(Bj.w)*kw1<T>(p.kern, sj);
    if (VAR == 0) {
      dvdr = T(Aj.x)*dx + T(Aj.y)*dy + T(Aj.z)*dz;
      dvdr -= T(mA.x)*dx + T(mA.y)*dy + T(mA.z)*dz;
      dv -= mj*dvdr*wkerni;
    } else {
      const T dvx = T(Aj.x - mA.x), dvy = T(Aj.y - mA.y), dvz = T(Aj.z - mA.z);
      dvdr = dvx*dx + dvy*dy + dvz*dz;
    }
    T paux = pterm_i*wkerni + T(Bj.y)*wkernj;
    if (dvdr < T(0)) {
      const T winvrho = T(0.25)*(wkerni + wkernj)*(invrho_i + invrho_j);
      const T vsignal = snd_i + T(Bj.z) - beta*alpha*dvdr;
      paux -= alpha*vsignal*dvdr*winvrho;
      du -= T(0.5)*mj*alpha*vsignal*dvdr*dvdr*winvrho;
    }
    ax += mj*dx*paux; ay += mj*dy*paux; az += mj*dz*paux;
    if (VAR == 1) {
      const T pg = T(0.5)*(invhsqd_i*kwgrav<T>(p.kern, si) + zeta_i*wkerni +
                           invh_j*invh_j*kwgrav<T>(p.kern, sj) + T(a.zeta[j])*wkernj);
      gx += mj*dx*pg; gy += mj*dy*pg; gz += mj*dz*pg;
      gp += T(0.5)*mj*(invh_i*kwpot<T>(p.kern, si) + invh_j*kwpot<T>(p.kern, sj));
      dv -= mj*dvdr*wkerni;
    }
  }
  double dvd = (double) dv*mB.x;
  const double rinv = mB.x;
  double dud = (double) du - a.pressure[i]*dvd*rinv*a.invomega[i];
  a.ax[i] = (double) ax; a.ay[i] = (double) ay; a.az[i] = (double) az;
  a.dudt[i] = dud; a.div_v[i] = dvd;
  if (VAR == 1) {
    a.gx[i] = (double) gx; a.gy[i] = (double) gy; a.gz[i] = (double) gz; a.gpot[i] = (double) gp;
  }
}

// FP32 variant: pair terms in FP32 (hardware reciprocal and square root), partial sums over
// blocks of 16 pairs in registers, Kahan-compensated totals across blocks
template <int VAR>
__global__ __launch_bounds__(BS) void k_hydro_mixed(HydroArgs a) {
  constexpr int KB = 16;
  const int i = xcd_block(blockIdx.x, gridDim.x)*BS + threadIdx.x;
  wave_add_ll(&a.icount[1], i < a.N ? (long long) a.nn[i] : 0LL);
  if (i >= a.N) return;
  const int n = a.nn[i];
  if (n > a.cap) { atomicMax(&a.flags[1], n); return; }
  const PhysParams& p = a.p;
  const double4 me = a.posm[i], mA = a.A[i], mB = a.B[i];
  const float invh_i = (float) mA.w;
  const float invrho_i = (float) mB.x;
  const float pterm_i = (float) mB.y, snd_i = (float) mB.z, hf_i = (float) mB.w;
  const float alpha = (float) p.alpha_visc, beta = (float) p.beta_visc;
  const float invhsqd_i = invh_i*invh_i;
  const float zeta_i = VAR == 1 ? (float) a.zeta[i] : 0.f;
  KahanF ax, ay, az, du, dv, gx, gy, gz, gp;
  for (int k0 = 0; k0 < n; k0 += KB) {
    const int k1 = min(n, k0 + KB);
    float bax = 0.f, bay = 0.f, baz = 0.f, bdu = 0.f, bdv = 0.f;
    float bgx = 0.f, bgy = 0.f, bgz = 0.f, bgp = 0.f;
    for (int k = k0; k < k1; k++) {
      const int j = a.list[(size_t) k*a.N + i];
      const double4 pj = a.posm[j], Aj = a.A[j], Bj = a.B[j];
      double ddx = pj.x - me.x, ddy = pj.y - me.y, ddz = pj.z - me.z;
      nearest_image<double>(p, ddx, ddy, ddz);
      float dx = (float) ddx, dy = (float) ddy, dz = (float) ddz;
      const float mj = (float) pj.w;
      const float invh_j = (float) Aj.w;
      const float invrho_j = (float) Bj.x;
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
      const float wkernj = (float) Bj.w*kw1<float>(p.kern, sj);
      const float dvx = (float) (Aj.x - mA.x), dvy = (float) (Aj.y - mA.y), dvz = (float) (Aj.z - mA.z);
      const float dvdr = dvx*dx + dvy*dy + dvz*dz;
      bdv -= mj*dvdr*wkerni;
      float paux = pterm_i*wkerni + (float) Bj.y*wkernj;
      if (dvdr < 0.f) {
        const float winvrho = 0.25f*(wkerni + wkernj)*(invrho_i + invrho_j);
        const float vsignal = snd_i + (float) Bj.z - beta*alpha*dvdr;
        paux -= alpha*vsignal*dvdr*winvrho;
        bdu -= 0.5f*mj*alpha*vsignal*dvdr*dvdr*winvrho;
      }
      bax += mj*dx*paux; bay += mj*dy*paux; baz += mj*dz*paux;
      if (VAR == 1) {
        const float pg = 0.5f*(invhsqd_i*kwgrav<float>(p.kern, si) + zeta_i*wkerni +
                               invh_j*invh_j*kwgrav<float>(p.kern, sj) + (float) a.zeta[j]*wkernj);
        bgx += mj*dx*pg; bgy += mj*dy*pg; bgz += mj*dz*pg;
        bgp += 0.5f*mj*(invh_i*kwpot<float>(p.kern, si) + invh_j*kwpot<float>(p.kern, sj));
      }
    }
    ax.add(bax); ay.add(bay); az.add(baz); du.add(bdu); dv.add(bdv);
    if (VAR == 1) { gx.add(bgx); gy.add(bgy); gz.add(bgz); gp.add(bgp); }
  }
  const double dvd = (double) dv.s*mB.x;
  const double dud = (double) du.s - a.pressure[i]*dvd*mB.x*a.invomega[i];
  a.ax[i] = (double) ax.s; a.ay[i] = (double) ay.s; a.az[i] = (double) az.s;
  a.dudt[i] = dud; a.div_v[i] = dvd;
  if (VAR == 1) {
    a.gx[i] = (double) gx.s; a.gy[i] = (double) gy.s; a.gz[i] = (double) gz.s;
    a.gpot[i] = (double) gp.s;
  }
}

ListWalk list_walk(gdf_ctx* c, int mode) {
  ListWalk w;
  w.N = c->N; w.ngroups = c->ngroups; w.nleaf = c->phys.nleaf; w.mode = mode;
  w.grp_node = c->grp_node; w.grp_first = c->grp_first; w.grp_count = c->grp_count;
  w.first = c->first; w.last = c->last; w.left = c->left; w.right = c->right;
  w.geo_c = c->geo_c; w.geo_h = c->geo_h;
  w.x = c->pf[F_X]; w.y = c->pf[F_Y]; w.z = c->pf[F_Z];
  if (mode == 0) {
    w.cap = c->cap_g; w.list = c->list_g; w.nn = c->nn_g;
    w.hcand = c->pf[F_H]; w.htarget = c->df[D_HSEARCH];
  } else {
    w.cap = c->cap_s; w.list = c->list_s; w.nn = c->nn_s;
    w.hcand = c->pf[F_H]; w.htarget = c->pf[F_H];
  }
  w.mixed = c->cfg.precision == GDF_PREC_MIXED;
  w.stacks = c->stacks; w.counter = c->counters + 4; w.flags = c->counters;
  w.p = c->phys;
  return w;
}

int read_flags(gdf_ctx* c) {
  GDF_HIP(c, hipMemcpyAsync(c->h_counters, c->counters, sizeof(int)*8, hipMemcpyDeviceToHost,
                            c->stream));
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  return GDF_OK;
}

int launch_list_walk(gdf_ctx* c, int mode, int nopbc) {
  ListWalk w = list_walk(c, mode);
  if (nopbc) w.p.periodic = 0;
  k_zero_ints<<<1, BS, 0, c->stream>>>(8, c->counters + 4);
  const int grid = std::min(c->ngroups, kWalkSlots);
  const int m = mode == 0 ? 0 : (w.mixed ? 2 : 1);
  hipStream_t st = c->stream;
  if (w.p.periodic) {
    if (m == 0) k_build_lists<0, true><<<grid, kWave, 0, st>>>(w);
    else if (m == 1) k_build_lists<1, true><<<grid, kWave, 0, st>>>(w);
    else k_build_lists<2, true><<<grid, kWave, 0, st>>>(w);
  } else {
    if (m == 0) k_build_lists<0, false><<<grid, kWave, 0, st>>>(w);
    else if (m == 1) k_build_lists<1, false><<<grid, kWave, 0, st>>>(w);
    else k_build_lists<2, false><<<grid, kWave, 0, st>>>(w);
  }
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int grow_list(gdf_ctx* c, int which, int need) {
  int& cap = which == 0 ? c->cap_g : c->cap_s;
  int*& list = which == 0 ? c->list_g : c->list_s;
  int ncap = std::max(need + need/4 + 8, cap*2);
  GDF_HIP(c, hipStreamSynchronize(c->stream));
  GDF_HIP(c, hipFree(list));
  list = nullptr;
  if (hipMalloc(&list, sizeof(int)*(size_t) ncap*(size_t) c->Ncap) != hipSuccess)
    return fail(c, GDF_ERR_NOMEM, "neighbour list allocation failed");
  cap = ncap;
  return GDF_OK;
}

}  // namespace

int sph_alloc_lists(gdf_ctx* c) {
  // initial capacity from the expected neighbour count of the kernel (Sph.cpp:106)
  const double kh = c->phys.kern.range*c->cfg.h_fac;
  const int ngather = (int) (4.0*3.14159265358979*kh*kh*kh/3.0);
  c->cap_g = std::max(64, (int) (ngather*1.8) + 32);
  c->cap_s = std::max(64, (int) (ngather*1.6) + 32);
  GDF_HIP(c, hipMalloc(&c->list_g, sizeof(int)*(size_t) c->cap_g*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->list_s, sizeof(int)*(size_t) c->cap_s*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->nn_g, sizeof(int)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->nn_s, sizeof(int)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->status, sizeof(int)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->hrecA, sizeof(double4)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->hrecB, sizeof(double4)*(size_t) c->Ncap));
  GDF_HIP(c, hipMalloc(&c->stacks, sizeof(int)*(size_t) kWalkSlots*kStackCap));
  return GDF_OK;
}

int sph_density(gdf_ctx* c, int nopbc) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (!c->tree_valid) return fail(c, GDF_ERR_INVALID, "density pass before tree build");
  k_init_density<<<nblocks(N, BS), BS, 0, s>>>(N, c->pf[F_H], c->df[D_HSEARCH], c->status);
  k_zero_ints<<<1, BS, 0, s>>>(4, c->counters);
  DensArgs d;
  d.N = N; d.nopbc = nopbc;
  d.posm = c->posm;
  d.h = c->pf[F_H]; d.hsearch = c->df[D_HSEARCH]; d.rho = c->df[D_RHO];
  d.invomega = c->df[D_INVOMEGA]; d.zeta = c->df[D_ZETA]; d.hfactor = c->df[D_HFACTOR];
  d.u = c->pf[F_U]; d.sound = c->df[D_SOUND]; d.pressure = c->df[D_PRESSURE];
  d.status = c->status; d.flags = c->counters; d.icount = c->icount;
  d.p = c->phys;
  for (int round = 0; round < 64; round++) {
    GDF_TRY(launch_list_walk(c, 0, nopbc));
    d.list = c->list_g; d.nn = c->nn_g; d.cap = c->cap_g;
    if (c->cfg.precision == GDF_PREC_MIXED)
      k_density<float><<<nblocks(N, BS), BS, 0, s>>>(d);
    else
      k_density<double><<<nblocks(N, BS), BS, 0, s>>>(d);
    GDF_HIP(c, hipGetLastError());
    GDF_TRY(read_flags(c));
    const int* f = c->h_counters;
    if (f[2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in gather pass");
    if (f[3]) return fail(c, GDF_ERR_NOCONV, "Problem with convergence of h-rho iteration");
    if (f[1] > 0) {
      GDF_TRY(grow_list(c, 0, f[1]));
      k_zero_ints<<<1, BS, 0, s>>>(4, c->counters);
      continue;  // status still marks every unfinished particle
    }
    if (f[0] == 0) {
      c->sg_valid = false;
      return tree_refit_h(c);
    }
    k_zero_ints<<<1, BS, 0, s>>>(4, c->counters);
  }
  return fail(c, GDF_ERR_CAPACITY, "h-rho neighbour search did not settle after 64 rounds");
}

int sph_build_sg(gdf_ctx* c) {
  for (int round = 0; round < 8; round++) {
    k_zero_ints<<<1, BS, 0, c->stream>>>(4, c->counters);
    GDF_TRY(launch_list_walk(c, 1, 0));
    GDF_TRY(read_flags(c));
    if (c->h_counters[2]) return fail(c, GDF_ERR_CAPACITY, "tree walk stack overflow in force pass");
    if (c->h_counters[1] > 0) { GDF_TRY(grow_list(c, 1, c->h_counters[1])); continue; }
    c->sg_valid = true;
    return GDF_OK;
  }
  return fail(c, GDF_ERR_CAPACITY, "scatter-gather list capacity");
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
  h.N = N; h.cap = c->cap_s; h.list = c->list_s; h.nn = c->nn_s;
  h.posm = c->posm; h.A = c->hrecA; h.B = c->hrecB;
  h.invomega = c->df[D_INVOMEGA]; h.pressure = c->df[D_PRESSURE]; h.zeta = c->df[D_ZETA];
  h.ax = c->df[D_AX]; h.ay = c->df[D_AY]; h.az = c->df[D_AZ];
  h.gx = c->df[D_ATX]; h.gy = c->df[D_ATY]; h.gz = c->df[D_ATZ]; h.gpot = c->df[D_GPOT];
  h.dudt = c->df[D_DUDT]; h.div_v = c->df[D_DIVV]; h.flags = c->counters; h.p = c->phys;
  h.icount = c->icount;
  const bool mixed = c->cfg.precision == GDF_PREC_MIXED;
  if (variant == 0) {
    if (mixed) k_hydro_mixed<0><<<nblocks(N, BS), BS, 0, s>>>(h);
    else k_hydro<double, 0><<<nblocks(N, BS), BS, 0, s>>>(h);
  } else {
    if (mixed) k_hydro_mixed<1><<<nblocks(N, BS), BS, 0, s>>>(h);
    else k_hydro<double, 1><<<nblocks(N, BS), BS, 0, s>>>(h);
  }
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

}  // namespace gdf
