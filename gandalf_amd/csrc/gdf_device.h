// Device-side helpers: smoothing kernels, periodic images, wave primitives.
//
// Kernel polynomials restate src/Headers/SmoothingKernel.h:132-408 (M4 and quintic, 3D) with
// the reference's own constants; powers are formed by multiplication.  Tabulated lookups follow
// SmoothingKernel.h:611-657 (nearest-lower entry of a 1000-entry table).
#pragma once

#include <hip/hip_runtime.h>
#include "gdf_internal.h"

namespace gdf {

// ---- fast reciprocal and square-root forms ----
// FP32: the hardware approximations (about 1 ulp), for the mixed-precision pair sums
__device__ __forceinline__ float rcp_f(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq_f(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float sqrt_f(float x) { return __builtin_amdgcn_sqrtf(x); }
// FP64 1/sqrt(x) for x > 0: hardware estimate (about 2^-23 relative) refined by two Newton
// steps y <- y + y (1 - x y^2)/2, which leaves about an ulp
__device__ __forceinline__ double rsq_d(double x) {
  double y = __builtin_amdgcn_rsq(x);
  #pragma unroll
  for (int it = 0; it < 2; it++) {
    const double e = fma(-(x*y), y, 1.0);
    y = fma(0.5*y, e, y);
  }
  return y;
}

enum { KW0 = 0, KW1, KWOMEGA, KWZETA, KWGRAV, KWPOT, KW0S2, KWOMEGAS2, KWZETAS2 };

template <typename T> __device__ __forceinline__ const T* ktab(const KernelDesc& k);
template <> __device__ __forceinline__ const double* ktab<double>(const KernelDesc& k) { return k.tab; }
template <> __device__ __forceinline__ const float* ktab<float>(const KernelDesc& k) { return k.tabf; }

// ---- M4 (SmoothingKernel.h:132-240) ----
template <typename T> __device__ __forceinline__ T m4_w0(T norm, T s) {
  if (s < T(1)) return norm*(T(1) - T(1.5)*s*s + T(0.75)*s*s*s);
  if (s < T(2)) { T q = T(2) - s; return T(0.25)*norm*(q*q*q); }
  return T(0);
}
template <typename T> __device__ __forceinline__ T m4_w1(T norm, T s) {
  if (s < T(1)) return norm*(T(-3)*s + T(2.25)*s*s);
  if (s < T(2)) return T(-0.75)*norm*(T(2) - s)*(T(2) - s);
  return T(0);
}
template <typename T> __device__ __forceinline__ T m4_womega(T norm, T s) {
  const T s3 = s*s*s;
  if (s < T(1)) return norm*(T(-3) + T(1.5)*T(5)*s*s - T(0.75)*T(6)*s3);
  if (s < T(2)) return norm*(T(-6) + T(3)*T(4)*s - T(1.5)*T(5)*s*s + T(0.25)*T(6)*s3);
  return T(0);
}
template <typename T> __device__ __forceinline__ T m4_wzeta(T s) {
  const T s2 = s*s, s4 = s2*s2, s5 = s4*s;
  if (s < T(1)) return T(1.4) - T(2)*s2 + T(1.5)*s4 - T(0.6)*s5;
  if (s < T(2)) return T(1.6) - T(4)*s2 + T(4)*s2*s - T(1.5)*s4 + T(0.2)*s5;
  return T(0);
}
template <typename T> __device__ __forceinline__ T m4_wgrav(T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2;
  if (s < T(1)) return T(1.333333333333333333333)*s - T(1.2)*s3 + T(0.5)*s4;
  if (s < T(2))
    return T(2.6666666666666666667)*s - T(3)*s2 + T(1.2)*s3 - T(0.166666666666666666667)*s4 -
           T(0.06666666666666666667)/s2;
  return T(1)/s2;
}
template <typename T> __device__ __forceinline__ T m4_wpot(T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s;
  if (s < T(1)) return T(1.4) - T(0.666666666666666666666666)*s2 + T(0.3)*s4 - T(0.1)*s5;
  if (s < T(2))
    return T(-1)/(T(15)*s) + T(1.6) - T(1.33333333333333333333333333)*s2 + s3 - T(0.3)*s4 +
           T(1.0/30.0)*s5;
  return T(1)/s;
}

// ---- quintic (SmoothingKernel.h:282-408) ----
template <typename T> __device__ __forceinline__ T q_w0(T norm, T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s;
  if (s < T(1)) return norm*(T(66) - T(60)*s2 + T(30)*s4 - T(10)*s5);
  if (s < T(2)) return norm*(T(51) + T(75)*s - T(210)*s2 + T(150)*s3 - T(45)*s4 + T(5)*s5);
  if (s < T(3)) return norm*(T(243) - T(405)*s + T(270)*s2 - T(90)*s3 + T(15)*s4 - s5);
  return T(0);
}
template <typename T> __device__ __forceinline__ T q_w1(T norm, T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2;
  if (s < T(1)) return norm*(T(-120)*s + T(120)*s3 - T(50)*s4);
  if (s < T(2)) return norm*(T(75) - T(420)*s + T(450)*s2 - T(180)*s3 + T(25)*s4);
  if (s < T(3)) return norm*(T(-405) + T(540)*s - T(270)*s2 + T(60)*s3 - T(5)*s4);
  return T(0);
}
template <typename T> __device__ __forceinline__ T q_womega(T norm, T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s;
  if (s < T(1)) return norm*(T(-198) + T(60)*T(5)*s2 - T(30)*T(7)*s4 + T(10)*T(8)*s5);
  if (s < T(2))
    return norm*(T(-153) - T(75)*T(4)*s + T(210)*T(5)*s2 - T(150)*T(6)*s3 + T(45)*T(7)*s4 -
                 T(5)*T(8)*s5);
  if (s < T(3))
    return norm*(T(-729) + T(405)*T(4)*s - T(270)*T(5)*s2 + T(90)*T(6)*s3 - T(15)*T(7)*s4 +
                 T(8)*s5);
  return T(0);
}
template <typename T> __device__ __forceinline__ T q_wzeta(T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s, s6 = s3*s3, s7 = s6*s;
  if (s < T(1)) return T(33)*s2 - T(15)*s4 + T(5)*s6 - T(1.42857142857)*s7 - T(34.14285714);
  if (s < T(2))
    return T(25.5)*s2 + T(25)*s3 - T(52.5)*s4 + T(30)*s5 - T(7.5)*s6 + T(0.7142857143)*s7 -
           T(33.785714286);
  if (s < T(3))
    return T(121.5)*s2 - T(135)*s3 + T(67.5)*s4 - T(18)*s5 + T(2.5)*s6 - T(0.142857143)*s7 -
           T(52.07142857);
  return T(0);
}
template <typename T> __device__ __forceinline__ T q_wgrav(T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s, s6 = s3*s3;
  const T c = T(12.0/359.0);
  if (s < T(1)) return c*(T(22)*s - T(12)*s3 + T(30.0/7.0)*s5 - T(5.0/4.0)*s6);
  if (s < T(2))
    return c*(T(17)*s + T(75.0/4.0)*s2 - T(42)*s3 + T(25)*s4 - T(45.0/7.0)*s5 +
              T(5.0/8.0)*s6 + T(5.0/56.0)/s2);
  if (s < T(3))
    return c*(T(81)*s - T(405.0/4.0)*s2 + T(54)*s3 - T(15)*s4 + T(15.0/7.0)*s5 -
              T(1.0/8.0)*s6 - T(507.0/56.0)/s2);
  return T(1)/s2;
}
template <typename T> __device__ __forceinline__ T q_wpot(T s) {
  const T s2 = s*s, s3 = s2*s, s4 = s2*s2, s5 = s4*s, s6 = s3*s3, s7 = s6*s;
  const T c = T(12.0/359.0);
  if (s < T(1))
    return c*(T(-11)*s2 + T(3)*s4 - T(5.0/7.0)*s6 + T(5.0/28.0)*s7 + T(478.0/14.0));
  if (s < T(2))
    return c*(T(-17.0/2.0)*s2 - T(25.0/4.0)*s3 + T(21.0/2.0)*s4 - T(5)*s5 + T(15.0/14.0)*s6 -
              T(5.0/56.0)*s7 + T(473.0/14.0) + T(5.0/56.0)/s);
  if (s < T(3))
    return c*(T(-81.0/2.0)*s2 + T(135.0/4.0)*s3 - T(27.0/2.0)*s4 + T(3)*s5 - T(5.0/14.0)*s6 +
              T(1.0/56.0)*s7 + T(729.0/14.0) - T(507.0/56.0)/s);
  return T(1)/s;
}

// ---- dispatch; s is r/h (or (r/h)^2 for the _s2 forms) ----
template <typename T> __device__ __forceinline__ T tab_lin(const KernelDesc& k, int w, T s) {
  if (s >= T(k.range)) return T(0);
  return ktab<T>(k)[w*k.res + (int) (s*T(k.resinv))];
}
template <typename T> __device__ __forceinline__ T tab_sq(const KernelDesc& k, int w, T s2) {
  if (s2 >= T(k.rangesqd)) return T(0);
  return ktab<T>(k)[w*k.res + (int) (s2*T(k.resinvsqd))];
}

template <typename T> __device__ __forceinline__ T kw0(const KernelDesc& k, T s) {
  if (k.tabulated) return tab_lin<T>(k, KW0, s);
  return k.type == 0 ? m4_w0<T>(T(k.norm), s) : q_w0<T>(T(k.norm), s);
}
// ---- branch-free analytic forms for the pair loops ----
// W and dW/ds as truncated powers, i.e. the reference's piecewise polynomials summed without the
// cancellation of their expanded form: M4 W = (norm/4)[(2-s)^3_+ - 4(1-s)^3_+], quintic
// W = norm[(3-s)^5_+ - 6(2-s)^5_+ + 15(1-s)^5_+]; womega = -3 w0 - s w1 exactly for both.
template <typename T> __device__ __forceinline__ T tpos(T x) { return x > T(0) ? x : T(0); }

template <typename T>
__device__ __forceinline__ void kw01_an(int type, T norm, T s, T& w0, T& w1) {
  if (type == 0) {
    const T A = tpos(T(2) - s), B = tpos(T(1) - s);
    const T A2 = A*A, B2 = B*B;
    w0 = T(0.25)*norm*(A2*A - T(4)*(B2*B));
    w1 = T(-0.75)*norm*(A2 - T(4)*B2);
  } else {
    const T A = tpos(T(3) - s), B = tpos(T(2) - s), C = tpos(T(1) - s);
    const T A2 = A*A, B2 = B*B, C2 = C*C;
    const T A4 = A2*A2, B4 = B2*B2, C4 = C2*C2;
    w0 = norm*(A4*A - T(6)*(B4*B) + T(15)*(C4*C));
    w1 = T(-5)*norm*(A4 - T(6)*B4 + T(15)*C4);
  }
}

template <typename T> __device__ __forceinline__ T kw1(const KernelDesc& k, T s) {
  if (k.tabulated) return tab_lin<T>(k, KW1, s);
  const T norm = T(k.norm);
  if (k.type == 0) {
    const T A = tpos(T(2) - s), B = tpos(T(1) - s);
    return T(-0.75)*norm*(A*A - T(4)*(B*B));
  }
  const T A = tpos(T(3) - s), B = tpos(T(2) - s), C = tpos(T(1) - s);
  const T A2 = A*A, B2 = B*B, C2 = C*C;
  return T(-5)*norm*(A2*A2 - T(6)*(B2*B2) + T(15)*(C2*C2));
}

template <typename T> __device__ __forceinline__ T kwgrav(const KernelDesc& k, T s) {
  if (k.tabulated) {
    if (s >= T(k.range)) return T(1)/(s*s);
    return ktab<T>(k)[KWGRAV*k.res + (int) (s*T(k.resinv))];
  }
  return k.type == 0 ? m4_wgrav<T>(s) : q_wgrav<T>(s);
}
template <typename T> __device__ __forceinline__ T kwpot(const KernelDesc& k, T s) {
  if (k.tabulated) {
    if (s >= T(k.range)) return T(1)/s;
    return ktab<T>(k)[KWPOT*k.res + (int) (s*T(k.resinv))];
  }
  return k.type == 0 ? m4_wpot<T>(s) : q_wpot<T>(s);
}
// ---- branch-free gravity kernels for the FP32 pair sums ----
// With W written as truncated powers, W = norm sum_k c_k (a_k - s)^n_+ (M4: n = 3, a = 2, 1,
// c = 1, -4; quintic: n = 5, a = 3, 2, 1, c = 1, -6, 15), the enclosed mass M(s) and the
// outer potential integral P(s) = 4 pi int_s^inf u W du are polynomials in x_k = (a_k - s)_+:
//   M = 1 - sum_k c_k x_k^(n+1) (a_k^2/(n+1) - 2 a_k x_k/(n+2) + x_k^2/(n+3)) * C,
//   P = sum_k c_k x_k^(n+1) (a_k/(n+1) - x_k/(n+2)) * C      (C = 1 for M4, 1/30 quintic),
// and wgrav = M/s^2, wpot = M/s + P, wzeta = wpot - s wgrav = P.  These are the reference's
// piecewise polynomials (SmoothingKernel.h:132-408) summed without branches; the reference's
// quintic wgrav and wpot carry 360/359 inside the kernel (its 12/359 in place of 1/30) and its
// quintic wzeta is -30 P, both kept.  Below s = 1 the mass comes from the inner polynomial,
// which avoids the cancellation of 1 - (...) at small s.  Agreement with the reference's
// polynomials: ~1e-8 relative (wgrav), ~1e-12 (wpot), within the reference's own constants
// (wzeta) -- far below FP32 rounding; the FP64 path keeps the piecewise forms.
__device__ __forceinline__ void kgravpot_f(const KernelDesc& k, float s, float inv, float& wg,
                                           float& wp) {
  if (k.tabulated) { wg = kwgrav<float>(k, s); wp = kwpot<float>(k, s); return; }
  const float x1 = fmaxf(1.f - s, 0.f), x2 = fmaxf(2.f - s, 0.f);
  const float x1s = x1*x1, x2s = x2*x2, s2 = s*s;
  if (k.type == 0) {
    const float x1q = x1s*x1s, x2q = x2s*x2s;
    const float D = fmaf(-4.f*x1q, fmaf(x1, fmaf(x1, 1.f/6.f, -0.4f), 0.25f),
                         x2q*fmaf(x2, fmaf(x2, 1.f/6.f, -0.8f), 1.f));
    const float P = fmaf(-4.f*x1q, fmaf(x1, -0.2f, 0.25f), x2q*fmaf(x2, -0.2f, 0.5f));
    const float Min = s2*s*fmaf(s2, fmaf(s, 0.5f, -1.2f), 4.f/3.f);
    const float M = s < 1.f ? Min : 1.f - D;
    wg = M*inv*inv;
    wp = fmaf(M, inv, P);
    return;
  }
  const float x3 = fmaxf(3.f - s, 0.f), x3s = x3*x3;
  const float x1h = x1s*x1s*x1s, x2h = x2s*x2s*x2s, x3h = x3s*x3s*x3s;
  const float D = fmaf(15.f*x1h, fmaf(x1, fmaf(x1, 0.125f, -2.f/7.f), 1.f/6.f),
                  fmaf(-6.f*x2h, fmaf(x2, fmaf(x2, 0.125f, -4.f/7.f), 2.f/3.f),
                       x3h*fmaf(x3, fmaf(x3, 0.125f, -6.f/7.f), 1.5f)));
  const float P30 = fmaf(15.f*x1h, fmaf(x1, -1.f/7.f, 1.f/6.f),
                    fmaf(-6.f*x2h, fmaf(x2, -1.f/7.f, 1.f/3.f), x3h*fmaf(x3, -1.f/7.f, 0.5f)));
  const float Min = s2*s*fmaf(s2, fmaf(s2, fmaf(s, -1.25f, 30.f/7.f), -12.f), 22.f)*(1.f/30.f);
  const float M = s < 1.f ? Min : fmaf(D, -1.f/30.f, 1.f);
  const float f = s < 3.f ? 360.f/359.f : 1.f;
  wg = f*M*inv*inv;
  wp = f*fmaf(M, inv, P30*(1.f/30.f));
}

// wzeta of the analytic kernels in FP32 without branches: P(s) for M4, -30 P(s) for quintic
__device__ __forceinline__ float kwzeta_f(int type, float s) {
  const float x1 = fmaxf(1.f - s, 0.f), x2 = fmaxf(2.f - s, 0.f);
  const float x1s = x1*x1, x2s = x2*x2;
  if (type == 0)
    return fmaf(-4.f*x1s*x1s, fmaf(x1, -0.2f, 0.25f), x2s*x2s*fmaf(x2, -0.2f, 0.5f));
  const float x3 = fmaxf(3.f - s, 0.f), x3s = x3*x3;
  return -fmaf(15.f*x1s*x1s*x1s, fmaf(x1, -1.f/7.f, 1.f/6.f),
               fmaf(-6.f*x2s*x2s*x2s, fmaf(x2, -1.f/7.f, 1.f/3.f),
                    x3s*x3s*x3s*fmaf(x3, -1.f/7.f, 0.5f)));
}

// density triple w0, womega, wzeta at s^2 (GradhSph.cpp:197-204)
template <typename T>
__device__ __forceinline__ void kdens(const KernelDesc& k, T s2, T& w0, T& wom, T& wz) {
  if (k.tabulated) {
    if (s2 >= T(k.rangesqd)) { w0 = wom = wz = T(0); return; }
    const int idx = (int) (s2*T(k.resinvsqd));
    const T* t = ktab<T>(k);
    w0 = t[KW0S2*k.res + idx]; wom = t[KWOMEGAS2*k.res + idx]; wz = t[KWZETAS2*k.res + idx];
    return;
  }
  const T s = sizeof(T) == 4 ? (T) sqrt_f((float) s2) : (T) sqrt((double) s2);
  T w1;
  kw01_an<T>(k.type, T(k.norm), s, w0, w1);
  wom = T(-3)*w0 - s*w1;
  if (sizeof(T) == 4) wz = (T) kwzeta_f(k.type, (float) s);
  else wz = k.type == 0 ? m4_wzeta<T>(s) : q_wzeta<T>(s);
}

// wzeta at s^2 alone (the stars' zeta terms, GradhSph.cpp:289-303)
__device__ __forceinline__ double kwzeta_s2(const KernelDesc& k, double s2) {
  if (k.tabulated) {
    if (s2 >= k.rangesqd) return 0.0;
    return k.tab[KWZETAS2*k.res + (int) (s2*k.resinvsqd)];
  }
  const double s = sqrt(s2);
  return k.type == 0 ? m4_wzeta<double>(s) : q_wzeta<double>(s);
}

// ---- periodic nearest image (GhostNeighbours.hpp:119-138) ----
template <typename T>
__device__ __forceinline__ void nearest_image(const PhysParams& p, T& dx, T& dy, T& dz) {
  if (!p.periodic) return;
  if (dx > T(p.boxhalf[0])) dx -= T(p.boxsize[0]); else if (dx < -T(p.boxhalf[0])) dx += T(p.boxsize[0]);
  if (dy > T(p.boxhalf[1])) dy -= T(p.boxsize[1]); else if (dy < -T(p.boxhalf[1])) dy += T(p.boxsize[1]);
  if (dz > T(p.boxhalf[2])) dz -= T(p.boxsize[2]); else if (dz < -T(p.boxhalf[2])) dz += T(p.boxsize[2]);
}

// Ewald::CalculatePeriodicCorrection (src/Common/Ewald.cpp:500-650) for a fully periodic box:
// the periodic correction of mass m at separation d = r_source - r_target.  A separation outside
// the half box first swaps its Newtonian term for that of its nearest image (the reference's
// repair of a non-nearest separation); the acceleration is trilinear in the table, the
// potential the value at the lower corner node, signs restored per axis.
template <typename T>
__device__ __forceinline__ void ewald_corr(const PhysParams& p, T dx, T dy, T dz, T m, T& ax,
                                           T& ay, T& az, T& pot) {
  T d[3] = {dx, dy, dz};
  const T half[3] = {T(p.boxhalf[0]), T(p.boxhalf[1]), T(p.boxhalf[2])};
  if (fabs(d[0]) > half[0] || fabs(d[1]) > half[1] || fabs(d[2]) > half[2]) {
    T inv = T(1)/sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2] + T(1.0e-20));
    T inv3 = inv*inv*inv;
    ax -= m*d[0]*inv3; ay -= m*d[1]*inv3; az -= m*d[2]*inv3;
    pot -= m*inv;
    for (int k = 0; k < 3; k++) {
      if (d[k] > half[k]) d[k] -= T(2)*half[k];
      else if (d[k] < -half[k]) d[k] += T(2)*half[k];
    }
    inv = T(1)/sqrt(d[0]*d[0] + d[1]*d[1] + d[2]*d[2] + T(1.0e-20));
    inv3 = inv*inv*inv;
    ax += m*d[0]*inv3; ay += m*d[1]*inv3; az += m*d[2]*inv3;
    pot += m*inv;
  }
  int il[3];
  T w[3];
  for (int k = 0; k < 3; k++) {
    const T b = T(p.edI[k])*fabs(d[k]);
    // |d| <= L/2 keeps b in [0, eng - 2] already; the clamp keeps any input in the table
    il[k] = max(0, min((int) b, p.eng[k] - 2));
    w[k] = b - (T) il[k];
  }
  const int n0 = p.eng[0], n01 = p.eng[0]*p.eng[1];
  const int b0 = il[0] + il[1]*n0 + il[2]*n01;
  T g0 = T(0), g1 = T(0), g2 = T(0), p0;
  auto corner = [&](int off, T wt) {
    if (sizeof(T) == 4) {
      const float4 t = p.etabf[b0 + off];
      g0 += wt*(T) t.y; g1 += wt*(T) t.z; g2 += wt*(T) t.w;
    } else {
      const double4 t = p.etab[b0 + off];
      g0 += wt*(T) t.y; g1 += wt*(T) t.z; g2 += wt*(T) t.w;
    }
  };
  const T u0 = T(1) - w[0], u1 = T(1) - w[1], u2 = T(1) - w[2];
  corner(0, u2*u1*u0);             corner(1, u2*u1*w[0]);
  corner(n0, u2*w[1]*u0);          corner(n0 + 1, u2*w[1]*w[0]);
  corner(n01, w[2]*u1*u0);         corner(n01 + 1, w[2]*u1*w[0]);
  corner(n01 + n0, w[2]*w[1]*u0);  corner(n01 + n0 + 1, w[2]*w[1]*w[0]);
  if (sizeof(T) == 4) p0 = (T) p.etabf[b0].x;
  else p0 = (T) p.etab[b0].x;
  pot += m*p0;
  const T sx = d[0] > T(0) ? T(1) : (d[0] < T(0) ? T(-1) : T(0));
  const T sy = d[1] > T(0) ? T(1) : (d[1] < T(0) ? T(-1) : T(0));
  const T sz = d[2] > T(0) ? T(1) : (d[2] < T(0) ? T(-1) : T(0));
  ax += m*g0*sx; ay += m*g1*sy; az += m*g2*sz;
}

// squared distance with the reference's operation order and no fused multiply-add, so
// neighbour-set membership is decided exactly as the oracle decides it.  (HIP's __dadd_rn and
// __dmul_rn are plain operators that the compiler may still contract; the pragma may not.)
__device__ __forceinline__ double dot3_rn(double a, double b, double c) {
#pragma clang fp contract(off)
  return a*a + b*b + c*c;
}
__device__ __forceinline__ double range2_rn(double krsqd, double h) {
#pragma clang fp contract(off)
  return krsqd*h*h;
}

// FP32 squared distance of the mixed path with a fixed operation order (explicit fused
// multiply-adds): the scatter-gather walk and the gravity walk both classify a pair with it,
// so each pair lands in exactly one of the two sums
__device__ __forceinline__ float dot3_f(float a, float b, float c) {
  return fmaf(c, c, fmaf(b, b, a*a));
}

// Kernel-range box test of the gravity walk with a relative slack: a node whose box comes within
// (1 + slack) k max(h) of the group is resolved to particles, so no pair of the scatter-gather
// set (decided per pair in FP64 or FP32) can end up inside an accepted multipole.
constexpr double kGravNeibSlack = 1.0 + 1.0e-4;

// ---- wave primitives (wave64) ----
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ int lanes_below(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned) (mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned) mask, 0));
}
__device__ __forceinline__ int wave_incl_scan(int v) {
  const int lane = __lane_id();
  #pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    int o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}
__device__ __forceinline__ int wave_max_int(int v) {
  #pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ double wave_max_d(double v) {
  #pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  return v;
}
__device__ __forceinline__ int wave_bcast(int v, int src) { return __shfl(v, src, 64); }
// sum over the wavefront, one atomic add by lane 0; every lane of the wave must take part
__device__ __forceinline__ void wave_add_ll(long long* dst, long long v) {
  #pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (__lane_id() == 0 && v != 0) atomicAdd((unsigned long long*) dst, (unsigned long long) v);
}

// The same over a whole workgroup (blockDim.x a multiple of 64, at most 1024; every thread
// calls it): one atomic per workgroup.  Atomics on one address from every wavefront of a
// million-particle pass serialise at the memory side (about 0.2 ms for 16384 of them).
__device__ __forceinline__ void block_add_ll(long long* dst, long long v) {
  __shared__ long long part[16];
  #pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if (__lane_id() == 0) part[threadIdx.x/64] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long t = 0;
    for (int w = 0; w < (int) blockDim.x/64; w++) t += part[w];
    if (t != 0) atomicAdd((unsigned long long*) dst, (unsigned long long) t);
  }
}

// GradhSph::ComputeThermalProperties (GradhSph.cpp:335-347) with the adiabatic
// (AdiabaticEOS.cpp), isothermal (IsothermalEOS.cpp) and barotropic (BarotropicEOS.cpp) closures:
// u (replaced for the last two), sound speed and pressure from rho and u
__device__ __forceinline__ void eos_state(const PhysParams& p, double rho, double& u,
                                          double& snd, double& press) {
  if (p.eos == 1) { u = p.temp0/p.gammam1/p.mu_bar; snd = sqrt(p.gammam1*u); }
  else if (p.eos == 2) {
    u = p.temp0*(1.0 + pow(rho*p.invrho_bary, p.gammam1))/p.gammam1/p.mu_bar;
    snd = sqrt(p.gammam1*u);
  } else snd = sqrt(p.gamma*p.gammam1*u);
  press = p.gammam1*rho*u;
}

// Next work item for a persistent wavefront.  Every active lane takes part in the atomic (the
// first active lane adds one, the others zero) and the result is broadcast with readfirstlane,
// so the index is a scalar and the loop that consumes it has uniform control flow.  Fetching
// behind `if (lane == 0)` and broadcasting with a shuffle instead lets the compiler treat the
// loop exit as divergent and re-enter the loop body with lane 0 split off, where the shuffle
// then reads a stale index.
__device__ __forceinline__ int wave_next_item(int* counter) {
  const int first = lanes_below(__ballot(1)) == 0;
  const int old = atomicAdd(counter, first ? 1 : 0);
  return __builtin_amdgcn_readfirstlane(old);
}

// MI355X has 8 XCDs, each with its own L2, and workgroups are dispatched to them round-robin by
// workgroup id.  Work that neighbours other work in memory (Morton-ordered particles, walk
// groups) is therefore remapped so each XCD takes one contiguous eighth.
constexpr int kXcds = 8;

// logical block of physical block b in a grid of nb: XCD x = b % 8 gets the x-th contiguous
// run of blocks (a bijection for any nb)
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int x = b % kXcds, k = b / kXcds, q = nb / kXcds, r = nb % kXcds;
  return x*q + min(x, r) + k;
}

// Next walk group for a persistent wavefront: the groups are split into 8 contiguous ranges with
// one counter each (counters[0..7]); a wave starts on its own XCD's range and moves to the next
// range when that is exhausted.  `range` carries the current range between calls; returns
// ngroups when every range is exhausted.
__device__ __forceinline__ int wave_next_group(int* counters, int ngroups, int& range) {
  for (int t = 0; t < kXcds; t++) {
    const int r = (range + t) % kXcds;
    const int lo = (int) (((long long) ngroups*r)/kXcds);
    const int hi = (int) (((long long) ngroups*(r + 1))/kXcds);
    const int k = wave_next_item(&counters[r]);
    if (lo + k < hi) { range = r; return lo + k; }
  }
  return ngroups;
}

// ---- tree walks ----
// A walk's node stack: the first kLdsStack entries live in LDS, deeper ones in the wave's
// global stack region (kStackCap entries), so the common case never waits on global memory.
constexpr int kLdsStack = 512;

__device__ __forceinline__ int stack_get(const int* lstk, const int* gstk, int s) {
  return s < kLdsStack ? lstk[s] : gstk[s - kLdsStack];
}
__device__ __forceinline__ void stack_put(int* lstk, int* gstk, int s, int v) {
  if (s < kLdsStack) lstk[s] = v;
  else gstk[s - kLdsStack] = v;
}

// particle range and children of a node (unified index: leaf l is node N - 1 + l), loaded
// together with the node's box so that a walk step waits on global memory once
struct NodeLinks {
  int first, count, left, right;
};
__device__ __forceinline__ void load_links(int n, int N, const int4* __restrict__ lnk,
                                           NodeLinks& l) {
  if (n >= N - 1) {
    l.first = n - (N - 1); l.count = 1; l.left = -1; l.right = -1;
  } else {
    const int4 k = lnk[n];           // (first, last, left, right)
    l.first = k.x; l.count = k.y - k.x + 1; l.left = k.z; l.right = k.w;
  }
}

// Queue of walk groups for a grid of persistent wavefronts.  nq entries (all groups, or a
// compacted list of the groups with targets): when they fit the grid, wavefront b takes entry b
// and stops, with no atomics and the work spread over every XCD; otherwise the entries come from
// the per-XCD ranged queues of wave_next_group.  Returns nq when the wavefront is done.
struct GroupQueue {
  int nq;
  int range;
  bool once;
};
__device__ __forceinline__ GroupQueue group_queue(int nq) {
  GroupQueue q;
  q.nq = nq; q.range = blockIdx.x % kXcds; q.once = false;
  return q;
}
__device__ __forceinline__ int group_next(GroupQueue& q, int* counters) {
  if (q.nq <= (int) gridDim.x) {
    if (q.once) return q.nq;
    q.once = true;
    return (int) blockIdx.x < q.nq ? (int) blockIdx.x : q.nq;
  }
  return wave_next_group(counters, q.nq, q.range);
}

// Kahan-compensated accumulator for the mixed-precision pair sums
struct KahanF {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace gdf
