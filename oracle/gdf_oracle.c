/* CPU oracle for the GANDALF particle hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C (OpenMP) restatement of the reference algorithms.  It is the checker for the HIP
 * path (tests/, the smoke check) and the CPU baseline timed beside it in the benchmark; the
 * product path never calls it.  Reference citations are to gandalfcode/gandalf src/.
 *
 * Arithmetic is written in the same operation order as the reference so that, compiled
 * without FMA contraction (-ffp-contract=off), FP64 results track the reference to rounding.
 */
#include "gdf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Constants.h:61-90 (the reference's truncated literals are part of its results) */
#define ORC_PI        3.14159265358979
#define ORC_TWOPI     6.28318530717959   /* Constants.h:62 */
#define ORC_INVPI     0.31830988618379
#define ORC_ONETHIRD  0.33333333333333333333333
#define SMALL_NUMBER  1.0e-20
#define SMALL_NUMBER_DP 1.0e-50
#define BIG_NUMBER    9.9e20
#define INVNDIM       (1.0/3)

/* ------------------------------------------------------------------------------------------ */
/* Smoothing kernels                                                                           */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int type, tabulated, res;
  double kernrange, kernrangesqd, kernnorm;
  double resinvkernrange, resinvkernrangesqd;
  double *tab;   /* 9 tables of res entries, ORC_* order */
} okern;

/* M4 kernel, 3D (SmoothingKernel.h:132-240, M4Kernel.cpp:36-55) */
static double m4_w0(const okern* k, double s) {
  if (s < 1.0) return (k->kernnorm)*(1.0 - 1.5*s*s + 0.75*s*s*s);
  else if (s < 2.0) return 0.25*(k->kernnorm)*pow(2.0 - s, 3);
  return 0.0;
}
static double m4_w1(const okern* k, double s) {
  if (s < 1.0) return (k->kernnorm)*(-3.0*s + 2.25*s*s);
  else if (s < 2.0) return -0.75*(k->kernnorm)*(2.0 - s)*(2.0 - s);
  return 0.0;
}
static double m4_womega(const okern* k, double s) {
  const int ndim = 3;
  if (s < 1.0)
    return (k->kernnorm)*(-(ndim) + 1.5*((ndim) + 2.0)*s*s - 0.75*((ndim) + 3.0)*pow(s,3));
  else if (s < 2.0)
    return (k->kernnorm)*(-2.0*(ndim) + 3.0*((ndim) + 1.0)*s - 1.50*((ndim) + 2.0)*s*s +
                          0.25*((ndim) + 3.0)*pow(s,3));
  return 0.0;
}
static double m4_wzeta(const okern* k, double s) {
  (void) k;
  if (s < 1.0) return 1.4 - 2.0*s*s + 1.5*pow(s,4) - 0.6*pow(s,5);
  else if (s < 2.0) return 1.6 - 4.0*s*s + 4.0*pow(s,3) - 1.5*pow(s,4) + 0.2*pow(s,5);
  return 0.0;
}
static double m4_wgrav(const okern* k, double s) {
  (void) k;
  if (s < 1.0) return 1.333333333333333333333*s - 1.2*pow(s,3) + 0.5*pow(s,4);
  else if (s < 2.0)
    return 2.6666666666666666667*s - 3.0*s*s + 1.2*pow(s,3) -
      0.166666666666666666667*pow(s,4) - 0.06666666666666666667/(s*s);
  return 1.0/(s*s);
}
static double m4_wpot(const okern* k, double s) {
  (void) k;
  if (s < 1.0) return 1.4 - 0.666666666666666666666666*s*s + 0.3*pow(s,4) - 0.1*pow(s,5);
  else if (s < 2.0)
    return -1.0/(15.0*s) + 1.6 - 1.33333333333333333333333333*s*s + pow(s,3) - 0.3*pow(s,4) +
      (1.0/30.0)*pow(s,5);
  return 1.0/s;
}

/* quintic kernel, 3D (SmoothingKernel.h:282-408, QuinticKernel.cpp:36-54) */
static double q_w0(const okern* k, double s) {
  if (s < 1.0) return (k->kernnorm)*(66.0 - 60.0*s*s + 30.0*pow(s,4) - 10.0*pow(s,5));
  else if (s < 2.0)
    return (k->kernnorm)*(51.0 + 75.0*s - 210.0*s*s + 150.0*pow(s,3) - 45.0*pow(s,4) +
                          5.0*pow(s,5));
  else if (s < 3.0)
    return (k->kernnorm)*(243.0 - 405*s + 270.0*s*s - 90.0*pow(s,3) + 15.0*pow(s,4) - pow(s,5));
  return 0.0;
}
static double q_w1(const okern* k, double s) {
  if (s < 1.0) return (k->kernnorm)*(-120.0*s + 120.0*pow(s,3) - 50.0*pow(s,4));
  else if (s < 2.0)
    return (k->kernnorm)*(75.0 - 420.0*s + 450.0*s*s - 180.0*pow(s,3) + 25.0*pow(s,4));
  else if (s < 3.0)
    return (k->kernnorm)*(-405.0 + 540.0*s - 270.0*s*s + 60.0*pow(s,3) - 5.0*pow(s,4));
  return 0.0;
}
static double q_womega(const okern* k, double s) {
  const int ndim = 3;
  if (s < 1.0)
    return (k->kernnorm)*(-66.0*(ndim) + 60.0*((ndim) + 2.0)*s*s -
                          30.0*((ndim) + 4.0)*pow(s,4) + 10.0*((ndim) + 5.0)*pow(s,5));
  else if (s < 2.0)
    return (k->kernnorm)*(-51.0*(ndim) - 75.0*((ndim) + 1.0)*s + 210.0*((ndim) + 2.0)*s*s -
                          150.0*((ndim) + 3.0)*pow(s,3) + 45.0*((ndim) + 4.0)*pow(s,4) -
                          5.0*((ndim) + 5.0)*pow(s,5));
  else if (s < 3.0)
    return (k->kernnorm)*(-243.0*(ndim) + 405.0*((ndim) + 1.0)*s - 270.0*((ndim) + 2.0)*s*s +
                          90.0*((ndim) + 3.0)*pow(s,3) - 15.0*((ndim) + 4.0)*pow(s,4) +
                          ((ndim) + 5.0)*pow(s,5));
  return 0.0;
}
static double q_wzeta(const okern* k, double s) {
  (void) k;
  if (s < 1.0)
    return 33.0*s*s - 15.0*pow(s,4) + 5.0*pow(s,6) - 1.42857142857*pow(s,7) - 34.14285714;
  else if (s < 2.0)
    return 25.5*s*s + 25.0*pow(s,3) - 52.5*pow(s,4) + 30.0*pow(s,5) - 7.5*pow(s,6) +
      0.7142857143*pow(s,7) - 33.785714286;
  else if (s < 3.0)
    return 121.5*s*s - 135.0*pow(s,3) + 67.5*pow(s,4) - 18.0*pow(s,5) + 2.5*pow(s,6) -
      0.142857143*pow(s,7) - 52.07142857;
  return 0.0;
}
static double q_wgrav(const okern* k, double s) {
  (void) k;
  if (s < 1.0)
    return (12.0/359.0)*(22.0*s - 12.0*pow(s,3) + (30.0/7.0)*pow(s,5) - (5.0/4.0)*pow(s,6));
  else if (s < 2.0)
    return (12.0/359.0)*(17.0*s + (75.0/4.0)*s*s - 42.0*pow(s,3) + 25.0*pow(s,4) -
                         (45.0/7.0)*pow(s,5) + (5.0/8.0)*pow(s,6) + (5.0/56.0)/(s*s));
  else if (s < 3.0)
    return (12.0/359.0)*(81.0*s - (405.0/4.0)*pow(s,2) + 54.0*pow(s,3) - 15.0*pow(s,4) +
                         (15.0/7.0)*pow(s,5) - (1.0/8.0)*pow(s,6) - (507.0/56.0)/(s*s));
  return 1.0/(s*s);
}
static double q_wpot(const okern* k, double s) {
  (void) k;
  if (s < 1.0)
    return (12.0/359.0)*(-11.0*s*s + 3.0*pow(s,4) - (5.0/7.0)*pow(s,6) + (5.0/28.0)*pow(s,7) +
                         (478.0/14.0));
  else if (s < 2.0)
    return (12.0/359.0)*(-(17.0/2.0)*s*s - (25.0/4.0)*pow(s,3) + (21.0/2.0)*pow(s,4) -
                         5.0*pow(s,5) + (15.0/14.0)*pow(s,6) - (5.0/56.0)*pow(s,7) +
                         (473.0/14.0) + (5.0/56.0)/s);
  else if (s < 3.0)
    return (12.0/359.0)*(-(81.0/2.0)*s*s + (135.0/4.0)*pow(s,3) - (27.0/2.0)*pow(s,4) +
                         3.0*pow(s,5) - (5.0/14.0)*pow(s,6) + (1.0/56.0)*pow(s,7) +
                         (729.0/14.0) - (507.0/56.0)/s);
  return 1.0/s;
}

typedef double (*kfun)(const okern*, double);
static kfun analytic_fun(int type, int which) {
  static const kfun m4[6] = {m4_w0, m4_w1, m4_womega, m4_wzeta, m4_wgrav, m4_wpot};
  static const kfun qu[6] = {q_w0, q_w1, q_womega, q_wzeta, q_wgrav, q_wpot};
  return type == 0 ? m4[which] : qu[which];
}

static void kern_init(okern* k, const orc_params* p) {
  memset(k, 0, sizeof(*k));
  k->type = p->kernel;
  k->tabulated = p->tabulated;
  k->res = 1000;
  if (p->kernel == 0) {
    k->kernrange = 2.0; k->kernrangesqd = 4.0; k->kernnorm = ORC_INVPI;
  } else {
    k->kernrange = 3.0; k->kernrangesqd = 9.0; k->kernnorm = ORC_INVPI*(1/120.);
  }
  k->resinvkernrange = k->res/k->kernrange;
  k->resinvkernrangesqd = k->res/k->kernrangesqd;
  if (k->tabulated) {
    /* TabulatedKernel.cpp:58-104 and SmoothingKernel.h:583-604 */
    k->tab = (double*) malloc(sizeof(double)*9*k->res);
    const double step = k->kernrange/k->res;
    const double stepsq = k->kernrangesqd/k->res;
    for (int w=0; w<6; w++) {
      kfun f = analytic_fun(k->type, w);
      for (int i=0; i<k->res; i++) k->tab[w*k->res + i] = f(k, step*i);
    }
    for (int w=0; w<3; w++) {
      kfun f = analytic_fun(k->type, w == 0 ? ORC_W0 : (w == 1 ? ORC_WOMEGA : ORC_WZETA));
      for (int i=0; i<k->res; i++) k->tab[(6 + w)*k->res + i] = f(k, sqrt(stepsq*i));
    }
  }
}
static void kern_free(okern* k) { free(k->tab); k->tab = NULL; }

/* SmoothingKernel.h:611-657 lookups */
static inline double tab_lookup(const okern* k, int w, double s) {
  if (s >= k->kernrange) return 0.0;
  double indexf = s*k->resinvkernrange;
  int index = (int) indexf;
  return k->tab[w*k->res + index];
}
static inline double tab_lookup_sqd(const okern* k, int w, double s2) {
  if (s2 >= k->kernrangesqd) return 0.0;
  double indexf = s2*k->resinvkernrangesqd;
  int index = (int) indexf;
  return k->tab[w*k->res + index];
}

static inline double K_w0(const okern* k, double s) {
  return k->tabulated ? tab_lookup(k, ORC_W0, s) : analytic_fun(k->type, ORC_W0)(k, s);
}
static inline double K_w1(const okern* k, double s) {
  return k->tabulated ? tab_lookup(k, ORC_W1, s) : analytic_fun(k->type, ORC_W1)(k, s);
}
static inline double K_womega(const okern* k, double s) {
  return k->tabulated ? tab_lookup(k, ORC_WOMEGA, s) : analytic_fun(k->type, ORC_WOMEGA)(k, s);
}
static inline double K_wzeta(const okern* k, double s) {
  return k->tabulated ? tab_lookup(k, ORC_WZETA, s) : analytic_fun(k->type, ORC_WZETA)(k, s);
}
static inline double K_wgrav(const okern* k, double s) {
  if (k->tabulated) {
    if (s >= k->kernrange) return 1.0/(s*s);
    return k->tab[ORC_WGRAV*k->res + (int)(s*k->resinvkernrange)];
  }
  return analytic_fun(k->type, ORC_WGRAV)(k, s);
}
static inline double K_wpot(const okern* k, double s) {
  if (k->tabulated) {
    if (s >= k->kernrange) return 1.0/s;
    return k->tab[ORC_WPOT*k->res + (int)(s*k->resinvkernrange)];
  }
  return analytic_fun(k->type, ORC_WPOT)(k, s);
}
static inline double K_w0_s2(const okern* k, double s2) {
  return k->tabulated ? tab_lookup_sqd(k, 6, s2) : analytic_fun(k->type, ORC_W0)(k, sqrt(s2));
}
static inline double K_womega_s2(const okern* k, double s2) {
  return k->tabulated ? tab_lookup_sqd(k, 7, s2)
                      : analytic_fun(k->type, ORC_WOMEGA)(k, sqrt(s2));
}
static inline double K_wzeta_s2(const okern* k, double s2) {
  return k->tabulated ? tab_lookup_sqd(k, 8, s2)
                      : analytic_fun(k->type, ORC_WZETA)(k, sqrt(s2));
}

double orc_kernel_eval(const orc_params* p, int which, double s) {
  okern k; kern_init(&k, p);
  double r = 0.0;
  switch (which) {
    case ORC_W0: r = K_w0(&k, s); break;
    case ORC_W1: r = K_w1(&k, s); break;
    case ORC_WOMEGA: r = K_womega(&k, s); break;
    case ORC_WZETA: r = K_wzeta(&k, s); break;
    case ORC_WGRAV: r = K_wgrav(&k, s); break;
    case ORC_WPOT: r = K_wpot(&k, s); break;
    case ORC_W0_S2: r = K_w0_s2(&k, s); break;
    case ORC_WOMEGA_S2: r = K_womega_s2(&k, s); break;
    case ORC_WZETA_S2: r = K_wzeta_s2(&k, s); break;
  }
  kern_free(&k);
  return r;
}

double orc_kernel_range(const orc_params* p) { return p->kernel == 0 ? 2.0 : 3.0; }

/* Sph.cpp:76-113, 3D branch: Ngather and h_guess computed with powf as the reference does */
double orc_h_guess(const orc_params* p, int N, const double* r) {
  double rmin[3] = {BIG_NUMBER, BIG_NUMBER, BIG_NUMBER};
  double rmax[3] = {-BIG_NUMBER, -BIG_NUMBER, -BIG_NUMBER};
  for (int i=0; i<N; i++)
    for (int k=0; k<3; k++) {
      if (r[3*i+k] < rmin[k]) rmin[k] = r[3*i+k];
      if (r[3*i+k] > rmax[k]) rmax[k] = r[3*i+k];
    }
  const double kernrange = orc_kernel_range(p);
  int Ngather = (int) (4.0*ORC_PI*pow(kernrange*p->h_fac,3)/3.0);
  double volume = (rmax[0] - rmin[0])*(rmax[1] - rmin[1])*(rmax[2] - rmin[2]);
  double h_guess = powf((3.0*volume*(double) Ngather)/(32.0*ORC_PI*(double) N), ORC_ONETHIRD);
  return h_guess;
}

/* ------------------------------------------------------------------------------------------ */
/* Periodic geometry (GhostNeighbours.hpp:119-138 NearestPeriodicVector)                      */
/* ------------------------------------------------------------------------------------------ */
typedef struct { int periodic; double size[3], half[3]; } obox;
static void box_init(obox* b, const orc_params* p, int nopbc) {
  b->periodic = p->periodic && !nopbc;
  for (int k=0; k<3; k++) {
    b->size[k] = p->boxmax[k] - p->boxmin[k];
    b->half[k] = 0.5*b->size[k];
  }
}
static inline void nearest(const obox* b, double* dr) {
  if (!b->periodic) return;
  for (int k=0; k<3; k++) {
    if (dr[k] > b->half[k]) dr[k] -= b->size[k];
    else if (dr[k] < -b->half[k]) dr[k] += b->size[k];
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Uniform cell grid for exact range queries                                                   */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
  int n[3];
  double lo[3], cell[3];
  int *start, *idx;  /* counting-sorted particle ids */
  int periodic;
} ogrid;

static void grid_build(ogrid* g, const obox* b, const orc_params* p, int N, const double* r,
                       double rsearch) {
  double lo[3], hi[3];
  for (int k=0; k<3; k++) { lo[k] = 1e300; hi[k] = -1e300; }
  for (int i=0; i<N; i++)
    for (int k=0; k<3; k++) {
      if (r[3*i+k] < lo[k]) lo[k] = r[3*i+k];
      if (r[3*i+k] > hi[k]) hi[k] = r[3*i+k];
    }
  g->periodic = b->periodic;
  long total = 1;
  for (int k=0; k<3; k++) {
    if (b->periodic) { lo[k] = p->boxmin[k]; hi[k] = p->boxmax[k]; }
    double ext = hi[k] - lo[k];
    int n = rsearch > 0.0 ? (int) floor(ext/rsearch) : 1;
    if (n < 1) n = 1;
    if (n > 128) n = 128;
    g->n[k] = n;
    g->lo[k] = lo[k];
    g->cell[k] = ext > 0.0 ? ext/n : 1.0;
    total *= n;
  }
  g->start = (int*) calloc(total + 1, sizeof(int));
  g->idx = (int*) malloc(sizeof(int)*(N > 0 ? N : 1));
  int* cid = (int*) malloc(sizeof(int)*(N > 0 ? N : 1));
  for (int i=0; i<N; i++) {
    int c[3];
    for (int k=0; k<3; k++) {
      int q = (int) floor((r[3*i+k] - g->lo[k])/g->cell[k]);
      if (q < 0) q = 0;
      if (q >= g->n[k]) q = g->n[k] - 1;
      c[k] = q;
    }
    cid[i] = (c[0]*g->n[1] + c[1])*g->n[2] + c[2];
    g->start[cid[i] + 1]++;
  }
  for (long c=0; c<total; c++) g->start[c+1] += g->start[c];
  int* fill = (int*) malloc(sizeof(int)*total);
  for (long c=0; c<total; c++) fill[c] = g->start[c];
  for (int i=0; i<N; i++) g->idx[fill[cid[i]]++] = i;
  free(fill);
  free(cid);
}
static void grid_free(ogrid* g) { free(g->start); free(g->idx); }

/* distinct neighbouring cell indices along one axis */
static int axis_cells(const ogrid* g, int k, int c, int* out) {
  int n = g->n[k];
  if (g->periodic && n <= 3) { for (int q=0; q<n; q++) out[q] = q; return n; }
  int m = 0;
  for (int d=-1; d<=1; d++) {
    int q = c + d;
    if (g->periodic) q = (q + n) % n;
    else if (q < 0 || q >= n) continue;
    out[m++] = q;
  }
  return m;
}

/* visit every candidate j of point x within the 27-cell stencil */
#define GRID_FOREACH(g, x, j, ...)                                                         \
  do {                                                                                      \
    int c_[3], ax_[3][3], na_[3];                                                           \
    for (int k_=0; k_<3; k_++) {                                                            \
      int q_ = (int) floor(((x)[k_] - (g)->lo[k_])/(g)->cell[k_]);                          \
      if (q_ < 0) q_ = 0;                                                                   \
      if (q_ >= (g)->n[k_]) q_ = (g)->n[k_] - 1;                                            \
      c_[k_] = q_;                                                                          \
      na_[k_] = axis_cells((g), k_, c_[k_], ax_[k_]);                                       \
    }                                                                                       \
    for (int a_=0; a_<na_[0]; a_++)                                                         \
      for (int b_=0; b_<na_[1]; b_++)                                                       \
        for (int e_=0; e_<na_[2]; e_++) {                                                   \
          int cc_ = (ax_[0][a_]*(g)->n[1] + ax_[1][b_])*(g)->n[2] + ax_[2][e_];             \
          for (int p_=(g)->start[cc_]; p_<(g)->start[cc_+1]; p_++) {                        \
            int j = (g)->idx[p_];                                                           \
            __VA_ARGS__                                                                     \
          }                                                                                 \
        }                                                                                   \
  } while (0)

static int cmp_int(const void* a, const void* b) {
  int x = *(const int*) a, y = *(const int*) b;
  return (x > y) - (x < y);
}

/* ------------------------------------------------------------------------------------------ */
/* Exact neighbour sets                                                                        */
/* ------------------------------------------------------------------------------------------ */
typedef struct { int* data; int n, cap; } ivec;
static void iv_push(ivec* v, int x) {
  if (v->n == v->cap) { v->cap = v->cap ? 2*v->cap : 64; v->data = realloc(v->data, sizeof(int)*v->cap); }
  v->data[v->n++] = x;
}

/* gather/search set of particle i for radius^2 hr2, strict or inclusive */
static void gather_one(const ogrid* g, const obox* b, const double* r, int i, double hr2,
                       int inclusive, ivec* out) {
  const double* xi = &r[3*i];
  out->n = 0;
  GRID_FOREACH(g, xi, j, {
    double dr[3] = {r[3*j] - xi[0], r[3*j+1] - xi[1], r[3*j+2] - xi[2]};
    nearest(b, dr);
    double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
    if (inclusive ? (drsqd <= hr2) : (drsqd < hr2)) iv_push(out, j);
  });
  qsort(out->data, out->n, sizeof(int), cmp_int);
}

int64_t orc_neighbour_sets(const orc_params* p, int N, const double* r, const double* h,
                           int mode, int64_t* offsets, int32_t* ids, int64_t cap) {
  obox b; box_init(&b, p, 0);
  const double kr = orc_kernel_range(p);
  const double krsqd = kr*kr;
  double hmax = 0.0;
  for (int i=0; i<N; i++) if (h[i] > hmax) hmax = h[i];
  ogrid g; grid_build(&g, &b, p, N, r, kr*hmax);
  ivec* lists = (ivec*) calloc(N, sizeof(ivec));
  #pragma omp parallel for schedule(dynamic, 64)
  for (int i=0; i<N; i++) {
    double hr2 = krsqd*h[i]*h[i];
    gather_one(&g, &b, r, i, hr2, mode == 2, &lists[i]);
  }
  if (mode == 1) {
    /* S_i = G_i u {j : i in G_j} (NeighbourManager.h:521: excluded iff r2 >= both ranges) */
    int* cnt = (int*) calloc(N, sizeof(int));
    for (int j=0; j<N; j++) for (int t=0; t<lists[j].n; t++) cnt[lists[j].data[t]]++;
    ivec* tr = (ivec*) calloc(N, sizeof(ivec));
    for (int i=0; i<N; i++) { tr[i].cap = cnt[i] ? cnt[i] : 1; tr[i].data = malloc(sizeof(int)*tr[i].cap); }
    for (int j=0; j<N; j++) for (int t=0; t<lists[j].n; t++) iv_push(&tr[lists[j].data[t]], j);
    #pragma omp parallel for schedule(dynamic, 64)
    for (int i=0; i<N; i++) {
      ivec u = {0, 0, 0};
      for (int t=0; t<lists[i].n; t++) iv_push(&u, lists[i].data[t]);
      for (int t=0; t<tr[i].n; t++) iv_push(&u, tr[i].data[t]);
      qsort(u.data, u.n, sizeof(int), cmp_int);
      int w = 0;
      for (int t=0; t<u.n; t++) if (w == 0 || u.data[w-1] != u.data[t]) u.data[w++] = u.data[t];
      u.n = w;
      free(lists[i].data);
      lists[i] = u;
    }
    for (int i=0; i<N; i++) free(tr[i].data);
    free(tr); free(cnt);
  }
  int64_t total = 0;
  offsets[0] = 0;
  for (int i=0; i<N; i++) { total += lists[i].n; offsets[i+1] = total; }
  if (total <= cap) {
    for (int i=0; i<N; i++)
      for (int t=0; t<lists[i].n; t++) ids[offsets[i] + t] = lists[i].data[t];
  }
  for (int i=0; i<N; i++) free(lists[i].data);
  free(lists);
  grid_free(&g);
  return total <= cap ? total : -total;
}

/* ------------------------------------------------------------------------------------------ */
/* EOS (EOS.h:156, AdiabaticEOS.cpp, IsothermalEOS.cpp, BarotropicEOS.cpp)                     */
/* ------------------------------------------------------------------------------------------ */
static void eos_thermal(const orc_params* p, double rho, double* u, double* sound,
                        double* pressure) {
  const double gammam1 = p->gamma - 1.0;
  if (p->eos == 1) {
    *u = p->temp0/gammam1/p->mu_bar;
    *sound = sqrt(gammam1*(*u));
  } else if (p->eos == 2) {
    *u = p->temp0*(1.0 + pow(rho*(1.0/p->rho_bary), gammam1))/gammam1/p->mu_bar;
    *sound = sqrt(gammam1*(*u));
  } else {
    *sound = sqrt(p->gamma*gammam1*(*u));
  }
  *pressure = gammam1*rho*(*u);
}

/* ------------------------------------------------------------------------------------------ */
/* grad-h h-rho iteration (GradhSph.cpp:142-326)                                               */
/* ------------------------------------------------------------------------------------------ */
static inline double h_rho_func(const orc_params* p, double m, double rho) {
  return p->h_fac*pow(m/rho, INVNDIM);
}
static inline double h_rho_deriv(double h, double m, double rho) {
  (void) m;
  return -INVNDIM*h/rho;
}

typedef struct { double h, rho, invomega, zeta, hfactor; } hres;

/* returns 1 ok, 0 = h grew beyond hmax (retry), -1 final h > hmax (accepted), -2 failure */
static int compute_h_one(const okern* kern, const orc_params* p, const double* xi, double m,
                         double h0, double hmax, int nn, const double* nx, const double* nm,
                         const orc_stars* st, hres* out) {
  int iteration = 0;
  const int iteration_max = 30;
  double h_lower_bound = 0.0, h_upper_bound = hmax;
  double h = h0, rho, invomega, zeta, hfactor, invh, invhsqd;
  do {
    iteration++;
    invh = 1.0/h;
    rho = 0.0; invomega = 0.0; zeta = 0.0;
    hfactor = pow(invh, 3);
    invhsqd = invh*invh;
    for (int j=0; j<nn; j++) {
      double dr0 = nx[3*j] - xi[0], dr1 = nx[3*j+1] - xi[1], dr2 = nx[3*j+2] - xi[2];
      double ssqd = invhsqd*(dr0*dr0 + dr1*dr1 + dr2*dr2);
      rho += nm[j]*K_w0_s2(kern, ssqd);
      invomega += nm[j]*invh*K_womega_s2(kern, ssqd);
      zeta += nm[j]*K_wzeta_s2(kern, ssqd);
    }
    rho *= hfactor;
    invomega *= hfactor;
    zeta *= invhsqd;
    if (rho > 0.0 && h > h_lower_bound &&
        fabs(h - h_rho_func(p, m, rho))*invh < p->h_converge) break;
    if (iteration < iteration_max) {
      h = h_rho_func(p, m, rho);
    } else if (iteration == iteration_max) {
      h = 0.5*(h_lower_bound + h_upper_bound);
    } else if (iteration < 5*iteration_max) {
      if (rho < SMALL_NUMBER || h > h_rho_func(p, m, rho)) h_upper_bound = h;
      else h_lower_bound = h;
      h = 0.5*(h_lower_bound + h_upper_bound);
    } else {
      return -2;
    }
    if (h > hmax) return 0;
  } while (h > h_lower_bound && h < h_upper_bound);

  h = fmax(h_rho_func(p, m, rho), h_lower_bound);
  invh = 1/h;
  out->hfactor = pow(invh, 4);
  invomega = 1.0 - h_rho_deriv(h, m, rho)*invomega;
  invomega = 1.0/invomega;
  /* the stars' conservative-gravity zeta terms (Hubber et al. 2013), GradhSph.cpp:282-303 */
  if (st && st->conservative) {
    for (int j=0; j<st->n; j++) {
      const double isq = st->softening ? pow(2.0/(h + st->h[j]), 2) : 4.0*invh*invh;
      const double dr0 = st->r[3*j] - xi[0], dr1 = st->r[3*j+1] - xi[1];
      const double dr2 = st->r[3*j+2] - xi[2];
      const double ssqd = (dr0*dr0 + dr1*dr1 + dr2*dr2)*isq;
      zeta += st->m[j]*isq*K_wzeta_s2(kern, ssqd);
    }
  }
  zeta = h_rho_deriv(h, m, rho)*zeta*invomega;
  out->h = h; out->rho = rho; out->invomega = invomega; out->zeta = zeta;
  return h <= hmax ? 1 : -1;
}

/* act (optional): only particles with act[i] != 0 are iterated, as UpdateAllSphProperties does
 * for the active particles of a block step (GradhSphTree.cpp:83-271) */
static int compute_h_masked(const orc_params* p, int N, const double* r, const double* m,
                            int nopbc, double* h, double* rho, double* invomega, double* zeta,
                            double* hfactor, double* u, double* sound, double* pressure,
                            const unsigned char* act, const orc_stars* st) {
  okern kern; kern_init(&kern, p);
  obox b; box_init(&b, p, nopbc);
  const double kr = kern.kernrange;
  double hmaxall = 0.0;
  for (int i=0; i<N; i++) if (h[i] > hmaxall) hmaxall = h[i];
  ogrid g; grid_build(&g, &b, p, N, r, kr*hmaxall*1.3);
  int fail = 0;
  #pragma omp parallel
  {
    int cap = 1024;
    double* nx = (double*) malloc(sizeof(double)*3*cap);
    double* nm = (double*) malloc(sizeof(double)*cap);
    #pragma omp for schedule(dynamic, 32)
    for (int i=0; i<N; i++) {
      if (act && !act[i]) continue;
      const double* xi = &r[3*i];
      double hmax = h[i];
      int ok;
      hres res;
      do {
        hmax = 1.05*hmax;
        const double hrangesqd = kern.kernrangesqd*hmax*hmax;
        int nn = 0;
        if (kr*hmax > kr*hmaxall*1.3) {
          /* search radius outgrew the grid cell: brute force */
          for (int j=0; j<N; j++) {
            double dr[3] = {r[3*j] - xi[0], r[3*j+1] - xi[1], r[3*j+2] - xi[2]};
            nearest(&b, dr);
            if (dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER <= hrangesqd) {
              if (nn == cap) { cap *= 2; nx = realloc(nx, sizeof(double)*3*cap); nm = realloc(nm, sizeof(double)*cap); }
              nx[3*nn] = xi[0] + dr[0]; nx[3*nn+1] = xi[1] + dr[1]; nx[3*nn+2] = xi[2] + dr[2];
              nm[nn++] = m[j];
            }
          }
        } else {
          /* candidates in ascending index order */
          ivec lst = {0, 0, 0};
          GRID_FOREACH(&g, xi, j, {
            double dr[3] = {r[3*j] - xi[0], r[3*j+1] - xi[1], r[3*j+2] - xi[2]};
            nearest(&b, dr);
            if (dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER <= hrangesqd) iv_push(&lst, j);
          });
          qsort(lst.data, lst.n, sizeof(int), cmp_int);
          if (lst.n > cap) { cap = lst.n; nx = realloc(nx, sizeof(double)*3*cap); nm = realloc(nm, sizeof(double)*cap); }
          for (int t=0; t<lst.n; t++) {
            int j = lst.data[t];
            double dr[3] = {r[3*j] - xi[0], r[3*j+1] - xi[1], r[3*j+2] - xi[2]};
            nearest(&b, dr);
            nx[3*nn] = xi[0] + dr[0]; nx[3*nn+1] = xi[1] + dr[1]; nx[3*nn+2] = xi[2] + dr[2];
            nm[nn++] = m[j];
          }
          free(lst.data);
        }
        ok = compute_h_one(&kern, p, xi, m[i], h[i], hmax, nn, nx, nm, st, &res);
      } while (ok == 0);
      if (ok == -2) {
        #pragma omp atomic write
        fail = 1;
        continue;
      }
      h[i] = res.h; rho[i] = res.rho; invomega[i] = res.invomega; zeta[i] = res.zeta;
      hfactor[i] = res.hfactor;
      eos_thermal(p, rho[i], &u[i], &sound[i], &pressure[i]);
    }
    free(nx); free(nm);
  }
  grid_free(&g);
  kern_free(&kern);
  return fail ? -1 : 0;
}

int orc_compute_h(const orc_params* p, int N, const double* r, const double* m, int nopbc,
                  double* h, double* rho, double* invomega, double* zeta, double* hfactor,
                  double* u, double* sound, double* pressure) {
  return compute_h_masked(p, N, r, m, nopbc, h, rho, invomega, zeta, hfactor, u, sound,
                          pressure, NULL, NULL);
}

int orc_compute_h_stars(const orc_params* p, int N, const double* r, const double* m,
                        int nopbc, const orc_stars* st, double* h, double* rho,
                        double* invomega, double* zeta, double* hfactor, double* u,
                        double* sound, double* pressure) {
  return compute_h_masked(p, N, r, m, nopbc, h, rho, invomega, zeta, hfactor, u, sound,
                          pressure, NULL, st);
}

/* ------------------------------------------------------------------------------------------ */
/* Stars beside the gas                                                                        */
/* ------------------------------------------------------------------------------------------ */
/* GradhSph::ComputeStarGravForces (GradhSph.cpp:699-743) for every gas particle: the stars'
 * kernel-softened pull over the mean smoothing length, added to a and gpot */
void orc_gas_from_stars(const orc_params* p, int N, const double* r, const double* h,
                        const orc_stars* st, double* a, double* gpot) {
  okern kern; kern_init(&kern, p);
  #pragma omp parallel for schedule(static)
  for (int i=0; i<N; i++) {
    for (int j=0; j<st->n; j++) {
      double dr[3];
      for (int k=0; k<3; k++) dr[k] = st->r[3*j+k] - r[3*i+k];
      const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER;
      const double drmag = sqrt(drsqd);
      const double invdrmag = 1.0/drmag;
      const double invhmean = 2.0/(h[i] + st->h[j]);
      const double paux = st->m[j]*invhmean*invhmean*K_wgrav(&kern, drmag*invhmean)*invdrmag;
      for (int k=0; k<3; k++) a[3*i+k] += paux*dr[k];
      gpot[i] += st->m[j]*invhmean*K_wpot(&kern, drmag*invhmean);
    }
  }
  kern_free(&kern);
}

/* HydroTree::UpdateAllStarGasForces (HydroTree.cpp:552-654) with the brute-force search's lists
 * (Tree.cpp:778-860 with one particle per cell): gas within (k h_s)/2 + (k/2) h_j of the star
 * in the softened sums of NbodyLeapfrogKDK::CalculateDirectHydroForces (NbodyLeapfrogKDK.cpp:
 * 178-200), the rest as point masses (:206-229); softened first, then direct, each in index
 * order.  Writes a and gpot of every star. */
void orc_stars_from_gas(const orc_params* p, const orc_stars* st, int N, const double* r,
                        const double* m, const double* h, double* a, double* gpot) {
  okern kern; kern_init(&kern, p);
  const double kr = kern.kernrange;
  for (int s=0; s<st->n; s++) {
    const double* rs = &st->r[3*s];
    double as[3] = {0.0, 0.0, 0.0}, pot = 0.0;
    for (int pass=0; pass<2; pass++) {
      for (int j=0; j<N; j++) {
        double dr[3];
        for (int k=0; k<3; k++) dr[k] = r[3*j+k] - rs[k];
        const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
        const double rn = 0.5*(kr*st->h[s]) + 0.5*kr*h[j];
        const int near = drsqd < pow(rn, 2);
        if (near != (pass == 0)) continue;
        const double drmag = sqrt(drsqd);
        const double invdrmag = 1.0/drmag;
        double paux;
        if (near) {
          const double invhmean = 2.0/(st->h[s] + h[j]);
          paux = m[j]*invhmean*invhmean*K_wgrav(&kern, drmag*invhmean)*invdrmag;
          pot += m[j]*invhmean*K_wpot(&kern, drmag*invhmean);
        } else {
          paux = m[j]*pow(invdrmag, 3);
          pot += m[j]*invdrmag;
        }
        for (int k=0; k<3; k++) as[k] += paux*dr[k];
      }
    }
    for (int k=0; k<3; k++) a[3*s+k] = as[k];
    gpot[s] = pot;
  }
  kern_free(&kern);
}

/* ------------------------------------------------------------------------------------------ */
/* Hydro forces                                                                                */
/* ------------------------------------------------------------------------------------------ */
static void scatter_gather_lists(const orc_params* p, const obox* b, int N, const double* r,
                                 const double* h, ivec** out) {
  int64_t* off = (int64_t*) malloc(sizeof(int64_t)*(N + 1));
  int64_t tot = orc_neighbour_sets(p, N, r, h, 1, off, NULL, 0);
  tot = tot < 0 ? -tot : tot;
  int32_t* ids = (int32_t*) malloc(sizeof(int32_t)*(tot > 0 ? tot : 1));
  orc_neighbour_sets(p, N, r, h, 1, off, ids, tot);
  ivec* l = (ivec*) calloc(N, sizeof(ivec));
  for (int i=0; i<N; i++) {
    l[i].n = l[i].cap = (int) (off[i+1] - off[i]);
    l[i].data = (int*) malloc(sizeof(int)*(l[i].n ? l[i].n : 1));
    for (int t=0; t<l[i].n; t++) l[i].data[t] = ids[off[i] + t];
  }
  free(ids); free(off);
  (void) b;
  *out = l;
}

/* the same sets for the rows with act[i] only (the others empty): a search of radius
 * k*max_j(h_j) around i filtered by the test above, r2 < k^2 h_i^2 or r2 < k^2 h_j^2; cheaper
 * than the transposition when few particles are active */
static void scatter_gather_lists_active(const orc_params* p, const obox* b, int N,
                                        const double* r, const double* h,
                                        const unsigned char* act, ivec** out) {
  const double kr = orc_kernel_range(p);
  const double krsqd = kr*kr;
  double hmax = 0.0;
  for (int i=0; i<N; i++) if (h[i] > hmax) hmax = h[i];
  ogrid g; grid_build(&g, b, p, N, r, kr*hmax);
  ivec* l = (ivec*) calloc(N, sizeof(ivec));
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i=0; i<N; i++) {
    if (!act[i]) continue;
    ivec cand = {0, 0, 0};
    gather_one(&g, b, r, i, krsqd*hmax*hmax, 1, &cand);
    const double* xi = &r[3*i];
    for (int t=0; t<cand.n; t++) {
      const int j = cand.data[t];
      double dr[3] = {r[3*j] - xi[0], r[3*j+1] - xi[1], r[3*j+2] - xi[2]};
      nearest(b, dr);
      const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
      if (drsqd < krsqd*h[i]*h[i] || drsqd < krsqd*h[j]*h[j]) iv_push(&l[i], j);
    }
    free(cand.data);
  }
  grid_free(&g);
  *out = l;
}

/* act (optional): forces of the active particles only; level/levelneib (optional): the
 * neighbour-level bookkeeping of GradhSph.cpp:445-446 merged as GradhSphTree.cpp:375-417:
 * for an active i and every j of its scatter-gather set (i itself included),
 * levelneib_i = max(levelneib_i, level_j) and levelneib_j = max(levelneib_j, level_i) */
static void hydro_forces_masked(const orc_params* p, int variant, int N, const double* r,
                                const double* v, const double* m, const double* h,
                                const double* rho, const double* invomega,
                                const double* pressure, const double* sound,
                                const double* hfactor, double* a, double* dudt, double* div_v,
                                const unsigned char* act, const int* level, int* levelneib) {
  okern kern; kern_init(&kern, p);
  obox b; box_init(&b, p, 0);
  ivec* lists;
  int nact = N;
  if (act) { nact = 0; for (int i=0; i<N; i++) nact += act[i] != 0; }
  if (act && nact < N/4) scatter_gather_lists_active(p, &b, N, r, h, act, &lists);
  else scatter_gather_lists(p, &b, N, r, h, &lists);
  const double alpha_visc = p->alpha_visc, beta_visc = p->beta_visc;
  const double gammam1 = p->gamma - 1.0;
  #pragma omp parallel for schedule(dynamic, 32)
  for (int i=0; i<N; i++) {
    if (act && !act[i]) continue;
    double ai[3] = {0.0, 0.0, 0.0}, du = 0.0, dv = 0.0;
    const double invh_i = 1/h[i];
    const double invrho_i = 1/rho[i];
    for (int t=0; t<lists[i].n; t++) {
      const int j = lists[i].data[t];
      const double invh_j = 1/h[j];
      const double invrho_j = 1/rho[j];
      double draux[3] = {r[3*j] - r[3*i], r[3*j+1] - r[3*i+1], r[3*j+2] - r[3*i+2]};
      nearest(&b, draux);
      double drmag, dvdr;
      if (variant == 0) {
        /* GradhSph.cpp:384-399 */
        drmag = sqrt(draux[0]*draux[0] + draux[1]*draux[1] + draux[2]*draux[2]);
        if (drmag > 0) for (int k=0; k<3; k++) draux[k] /= drmag;
      } else {
        /* GradhSph.cpp:498-505 */
        drmag = sqrt(draux[0]*draux[0] + draux[1]*draux[1] + draux[2]*draux[2] + SMALL_NUMBER);
        const double invdrmag = 1.0/drmag;
        for (int k=0; k<3; k++) draux[k] *= invdrmag;
      }
      const double wkerni = hfactor[i]*K_w1(&kern, drmag*invh_i);
      const double wkernj = hfactor[j]*K_w1(&kern, drmag*invh_j);
      if (variant == 0) {
        dvdr = v[3*j]*draux[0] + v[3*j+1]*draux[1] + v[3*j+2]*draux[2];
        dvdr -= v[3*i]*draux[0] + v[3*i+1]*draux[1] + v[3*i+2]*draux[2];
      } else {
        double dvv[3] = {v[3*j] - v[3*i], v[3*j+1] - v[3*i+1], v[3*j+2] - v[3*i+2]};
        dvdr = dvv[0]*draux[0] + dvv[1]*draux[1] + dvv[2]*draux[2];
      }
      if (variant == 0) dv -= m[j]*dvdr*wkerni;
      double paux = ((pressure[i]*invomega[i])/(rho[i]*rho[i]))*wkerni +
        ((pressure[j]*invomega[j])/(rho[j]*rho[j]))*wkernj;
      if (dvdr < 0.0) {
        const double winvrho = 0.25*(wkerni + wkernj)*(invrho_i + invrho_j);
        const double vsignal = sound[i] + sound[j] - beta_visc*alpha_visc*dvdr;
        paux -= alpha_visc*vsignal*dvdr*winvrho;
        du -= 0.5*m[j]*alpha_visc*vsignal*dvdr*dvdr*winvrho;
      }
      for (int k=0; k<3; k++) ai[k] += m[j]*draux[k]*paux;
      if (variant == 1) dv -= m[j]*dvdr*wkerni;
    }
    dv *= invrho_i;
    /* eos->Pressure(parti) = gammam1*rho*u; recomputed from the stored pressure's inputs */
    du -= pressure[i]*dv*invrho_i*invomega[i];
    (void) gammam1;
    for (int k=0; k<3; k++) a[3*i+k] = ai[k];
    dudt[i] = du;
    div_v[i] = dv;
  }
  if (level) {
    for (int i=0; i<N; i++) {
      if (act && !act[i]) continue;
      for (int t=0; t<lists[i].n; t++) {
        const int j = lists[i].data[t];
        if (level[j] > levelneib[i]) levelneib[i] = level[j];
        if (level[i] > levelneib[j]) levelneib[j] = level[i];
      }
    }
  }
  for (int i=0; i<N; i++) free(lists[i].data);
  free(lists);
  kern_free(&kern);
}

void orc_hydro_forces(const orc_params* p, int variant, int N, const double* r,
                      const double* v, const double* m, const double* h, const double* rho,
                      const double* invomega, const double* pressure, const double* sound,
                      const double* hfactor, double* a, double* dudt, double* div_v) {
  hydro_forces_masked(p, variant, N, r, v, m, h, rho, invomega, pressure, sound, hfactor, a,
                      dudt, div_v, NULL, NULL, NULL);
}

/* ------------------------------------------------------------------------------------------ */
/* Gravity pair terms                                                                          */
/* ------------------------------------------------------------------------------------------ */
/* one particle-particle term: kernel-softened if within either kernel range
 * (GradhSph.cpp:555-565 with drmag = sqrt(r2 + small)), else Newtonian (:671-686) */
static inline void grav_pair(const okern* kern, double krsqd, const double* dr_in, double mj,
                             double hi, double hj, double zi, double zj, double hfi, double hfj,
                             double* acc, double* pot) {
  double dr[3] = {dr_in[0], dr_in[1], dr_in[2]};
  const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
  if (drsqd < krsqd*hi*hi || drsqd < krsqd*hj*hj) {
    const double invh_i = 1/hi, invh_j = 1/hj;
    const double drmag = sqrt(drsqd + SMALL_NUMBER);
    const double invdrmag = 1.0/drmag;
    for (int k=0; k<3; k++) dr[k] *= invdrmag;
    const double wkerni = hfi*K_w1(kern, drmag*invh_i);
    const double wkernj = hfj*K_w1(kern, drmag*invh_j);
    const double paux = 0.5*(invh_i*invh_i*K_wgrav(kern, drmag*invh_i) + zi*wkerni +
                             invh_j*invh_j*K_wgrav(kern, drmag*invh_j) + zj*wkernj);
    for (int k=0; k<3; k++) acc[k] += mj*dr[k]*paux;
    *pot += 0.5*mj*(invh_i*K_wpot(kern, drmag*invh_i) + invh_j*K_wpot(kern, drmag*invh_j));
  } else {
    const double d2 = drsqd + SMALL_NUMBER;
    const double invdrmag = 1.0/sqrt(d2);
    const double invdr3 = invdrmag*invdrmag*invdrmag;
    for (int k=0; k<3; k++) acc[k] += mj*dr[k]*invdr3;
    *pot += mj*invdrmag;
  }
}

static void direct_gravity_masked(const orc_params* p, int N, const double* r,
                                  const double* m, const double* h, const double* zeta,
                                  const double* hfactor, double* atree, double* gpot,
                                  const unsigned char* act) {
  okern kern; kern_init(&kern, p);
  obox b; box_init(&b, p, 0);
  const double krsqd = kern.kernrangesqd;
  const double wpot0 = K_wpot(&kern, 0.0);
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i=0; i<N; i++) {
    if (act && !act[i]) continue;
    double acc[3] = {0.0, 0.0, 0.0};
    double pot = (m[i]/h[i])*wpot0;
    for (int j=0; j<N; j++) {
      double dr[3] = {r[3*j] - r[3*i], r[3*j+1] - r[3*i+1], r[3*j+2] - r[3*i+2]};
      nearest(&b, dr);
      grav_pair(&kern, krsqd, dr, m[j], h[i], h[j], zeta[i], zeta[j], hfactor[i], hfactor[j],
                acc, &pot);
    }
    for (int k=0; k<3; k++) atree[3*i+k] = acc[k];
    gpot[i] = pot;
  }
  kern_free(&kern);
}

void orc_direct_gravity(const orc_params* p, int N, const double* r, const double* m,
                        const double* h, const double* zeta, const double* hfactor,
                        double* atree, double* gpot) {
  direct_gravity_masked(p, N, r, m, h, zeta, hfactor, atree, gpot, NULL);
}

/* ------------------------------------------------------------------------------------------ */
/* Ewald periodic gravity                                                                      */
/* ------------------------------------------------------------------------------------------ */
/* Ewald::AccShort (Ewald.cpp:841-850) */
static double ewald_acc_short(double alpha, double proj, double r) {
  const double r2 = pow(r, 2);
  const double s1 = 2.0*alpha*proj*exp(-r2*pow(alpha, 2))/(sqrt(ORC_PI)*r2);
  const double s2 = proj*erfc(alpha*r)/pow(r, 3);
  return s1 + s2;
}

/* Ewald::AccLong3pPer (Ewald.cpp:1005-1013) */
static double ewald_acc_long(int l, double linv_acc, double dzeta, double c3, double c4) {
  return l*linv_acc*exp(-dzeta*c3)*sin(c4)/c3;
}

/* Ewald::Ewald (Ewald.cpp:42-420), fully periodic box (ewald_periodicity = 7): the correction
 * field, component-major as the reference stores it (potential, then a_x, a_y, a_z, each
 * Ng[0]*Ng[1]*Ng[2] nodes with x fastest) */
int orc_ewald_table(const double L[3], int nseries, int ngrid, double ewald_mult,
                    double efratio, double* field, int Ng[3], double dI[3]) {
  const double lx = L[0], ly = L[1], lz = L[2];
  const double linv = 1.0/lx;
  const double ratio_p1 = ly/lx, ratio_p2 = lz/lx;
  const int es_nrx = nseries, es_nry = (int) (1 + nseries/ratio_p1);
  const int es_nrz = (int) (1 + nseries/ratio_p2);
  const int es_nfx = nseries, es_nfy = (int) (1 + nseries*ratio_p1);
  const int es_nfz = (int) (1 + nseries*ratio_p2);
  const int es_radius2 = nseries*nseries;
  const double cr1 = 1.0, cr2 = pow(ratio_p1, 2), cr3 = pow(ratio_p2, 2);
  const double cf1 = 1.0, cf2 = 1.0/pow(ratio_p1, 2), cf3 = 1.0/pow(ratio_p2, 2);
  const double ratio_pinv1 = 1.0/ratio_p1, ratio_pinv2 = 1.0/ratio_p2;
  const double linv2 = linv*ratio_pinv1*ratio_pinv2/ORC_PI;
  const double linv_acc = 2.0*pow(linv, 2)*ratio_pinv1*ratio_pinv2;
  const double alpha = 2.0*linv*ewald_mult;
  const double dzeta = ORC_PI*ORC_PI*linv*linv/(alpha*alpha);
  double Lewald[3];
  int Ncells[3];
  for (int k=0; k<3; k++) {
    Lewald[k] = efratio*0.5*L[k];
    Ncells[k] = ngrid - 1;
    dI[k] = (Ncells[k] - 1)/Lewald[k];
    Ng[k] = Ncells[k] + 1;
  }
  if (!field) return 0;
  const int one = Ng[0]*Ng[1]*Ng[2];
  for (int i=0; i<4*one; i++) field[i] = 0.0;
  #pragma omp parallel for schedule(dynamic, 1)
  for (int k=0; k<Ng[2]; k++) {
    for (int j=0; j<Ng[1]; j++) {
      for (int i=0; i<Ng[0]; i++) {
        const int e = k*Ng[1]*Ng[0] + j*Ng[0] + i;
        const double x = i*Lewald[0]/(Ncells[0] - 1);
        const double y = j*Lewald[1]/(Ncells[1] - 1);
        const double z = k*Lewald[2]/(Ncells[2] - 1);
        for (int ni=-es_nrx; ni<=es_nrx; ni++)
          for (int nj=-es_nry; nj<=es_nry; nj++)
            for (int nk=-es_nrz; nk<=es_nrz; nk++) {
              if ((cr1*ni*ni + cr2*nj*nj + cr3*nk*nk) <= es_radius2) {
                double rrel[4];
                rrel[1] = x + ni*lx; rrel[2] = y + nj*ly; rrel[3] = z + nk*lz;
                const double rni2 = pow(rrel[1], 2) + pow(rrel[2], 2) + pow(rrel[3], 2);
                const double rni = sqrt(rni2);
                field[e] = field[e] + erfc(alpha*rni)/rni;
                for (int jj=1; jj<4; jj++) field[e + jj*one] += ewald_acc_short(alpha, rrel[jj], rni);
              }
            }
        for (int hi=-es_nfx; hi<=es_nfx; hi++)
          for (int hj=-es_nfy; hj<=es_nfy; hj++)
            for (int hk=-es_nfz; hk<=es_nfz; hk++) {
              if ((cf1*hi*hi + cf2*hj*hj + cf3*hk*hk) <= es_radius2 &&
                  (pow(hi, 2) + pow(hj, 2) + pow(hk, 2)) > 0) {
                const double c3 = pow(hi, 2) + pow(hj*ratio_pinv1, 2) + pow(hk*ratio_pinv2, 2);
                const double c4 = ORC_TWOPI*linv*(hi*x + hj*y*ratio_pinv1 + hk*z*ratio_pinv2);
                field[e] += linv2*exp(-dzeta*c3)*cos(c4)/c3;
                field[e + one] += ewald_acc_long(hi, linv_acc, dzeta, c3, c4);
                field[e + 2*one] += ratio_pinv1*(ewald_acc_long(hj, linv_acc, dzeta, c3, c4));
                field[e + 3*one] += ratio_pinv2*(ewald_acc_long(hk, linv_acc, dzeta, c3, c4));
              }
            }
        const double rni = sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2));
        field[e] -= 1.0/rni;
        field[e + one] -= x/pow(rni, 3);
        field[e + 2*one] -= y/pow(rni, 3);
        field[e + 3*one] -= z/pow(rni, 3);
      }
    }
  }
  field[0] = 0.0; field[one] = 0.0; field[2*one] = 0.0; field[3*one] = 0.0;
  return 0;
}

static inline double sgn_d(double v) { return (double) ((0.0 < v) - (v < 0.0)); }

/* Ewald::CalculatePeriodicCorrection (Ewald.cpp:500-650), fully periodic box */
void orc_ewald_corr(const orc_ewald* e, double m, const double* dr_in, double* acorr,
                    double* gpotcorr) {
  double dr[3] = {dr_in[0], dr_in[1], dr_in[2]};
  const double lper[3] = {e->L[0], e->L[1], e->L[2]};
  acorr[0] = acorr[1] = acorr[2] = 0.0;
  *gpotcorr = 0.0;
  if (fabs(dr[0]) > 0.5*lper[0] || fabs(dr[1]) > 0.5*lper[1] || fabs(dr[2]) > 0.5*lper[2]) {
    double inv_dr = 1/sqrt(dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER);
    double inv_dr3 = inv_dr*inv_dr*inv_dr;
    for (int k=0; k<3; k++) acorr[k] -= m*dr[k]*inv_dr3;
    *gpotcorr -= m*inv_dr;
    for (int k=0; k<3; k++) {
      const double size = 0.5*lper[k];
      if (dr[k] > size) dr[k] -= 2*size;
      else if (dr[k] < -size) dr[k] += 2*size;
    }
    inv_dr = 1/sqrt(dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER);
    inv_dr3 = inv_dr*inv_dr*inv_dr;
    for (int k=0; k<3; k++) acorr[k] += m*dr[k]*inv_dr3;
    *gpotcorr += m*inv_dr;
  }
  int il[3];
  double a[3];
  for (int k=0; k<3; k++) {
    const double b = e->dI[k]*fabs(dr[k]);
    il[k] = (int) b;
    a[k] = b - il[k];
  }
  double wf[8];
  wf[0] = (1.0 - a[2])*(1.0 - a[1])*(1.0 - a[0]);
  wf[1] = (1.0 - a[2])*(1.0 - a[1])*(      a[0]);
  wf[2] = (1.0 - a[2])*(      a[1])*(1.0 - a[0]);
  wf[3] = (1.0 - a[2])*(      a[1])*(      a[0]);
  wf[4] = (      a[2])*(1.0 - a[1])*(1.0 - a[0]);
  wf[5] = (      a[2])*(1.0 - a[1])*(      a[0]);
  wf[6] = (      a[2])*(      a[1])*(1.0 - a[0]);
  wf[7] = (      a[2])*(      a[1])*(      a[0]);
  const int n0 = e->Ng[0], n01 = e->Ng[0]*e->Ng[1], one = n01*e->Ng[2];
  const int i0 = il[0] + il[1]*n0 + il[2]*n01, i1 = i0 + n01;
  double g[4];
  for (int k=0; k<=3; k++) {
    const double* f = e->field + k*one;
    g[k] = wf[0]*f[i0] + wf[1]*f[i0 + 1] + wf[2]*f[i0 + n0] + wf[3]*f[i0 + n0 + 1] +
           wf[4]*f[i1] + wf[5]*f[i1 + 1] + wf[6]*f[i1 + n0] + wf[7]*f[i1 + n0 + 1];
  }
  *gpotcorr += m*e->field[i0];
  for (int k=0; k<3; k++) acorr[k] += m*g[k+1]*sgn_d(dr[k]);
}

/* the exact periodic gravity of a fully periodic box (GradhSphTree::UpdateAllSphForces with the
 * brute-force search and Ewald, GradhSphTree.cpp:515-578): every pair's kernel-softened or
 * direct term at its nearest image plus its Ewald correction, and the self potential */
void orc_direct_gravity_periodic(const orc_params* p, const orc_ewald* e, int N,
                                 const double* r, const double* m, const double* h,
                                 const double* zeta, const double* hfactor, double* atree,
                                 double* gpot) {
  okern kern; kern_init(&kern, p);
  obox b; box_init(&b, p, 0);
  const double krsqd = kern.kernrangesqd;
  const double wpot0 = K_wpot(&kern, 0.0);
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i=0; i<N; i++) {
    double acc[3] = {0.0, 0.0, 0.0};
    double pot = (m[i]/h[i])*wpot0;
    for (int j=0; j<N; j++) {
      double dr[3] = {r[3*j] - r[3*i], r[3*j+1] - r[3*i+1], r[3*j+2] - r[3*i+2]};
      nearest(&b, dr);
      grav_pair(&kern, krsqd, dr, m[j], h[i], h[j], zeta[i], zeta[j], hfactor[i], hfactor[j],
                acc, &pot);
      double ac[3], pc;
      orc_ewald_corr(e, m[j], dr, ac, &pc);
      for (int k=0; k<3; k++) acc[k] += ac[k];
      pot += pc;
    }
    for (int k=0; k<3; k++) atree[3*i+k] = acc[k];
    gpot[i] = pot;
  }
  kern_free(&kern);
}

void orc_direct_plummer(int N, const double* r, const double* m, double eps,
                        const int32_t* targets, int ntargets, double* a, double* pot) {
  const double eps2 = eps*eps;
  #pragma omp parallel for schedule(dynamic, 16)
  for (int t=0; t<ntargets; t++) {
    const int i = targets[t];
    double acc[3] = {0.0, 0.0, 0.0}, ph = 0.0;
    for (int j=0; j<N; j++) {
      if (j == i) continue;
      double dr[3] = {r[3*j] - r[3*i], r[3*j+1] - r[3*i+1], r[3*j+2] - r[3*i+2]};
      double d2 = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + eps2;
      double inv = 1.0/sqrt(d2);
      double inv3 = inv*inv*inv;
      for (int k=0; k<3; k++) acc[k] += m[j]*dr[k]*inv3;
      ph -= m[j]*inv;
    }
    for (int k=0; k<3; k++) a[3*t+k] = acc[k];
    pot[t] = ph;
  }
}


/* Nbody::CalculateDirectGravForces (src/Nbody/Nbody.cpp:233-293) */
void orc_nbody_direct(int N, const double* r, const double* v, const double* m, double* a,
                      double* adot, double* gpot) {
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i=0; i<N; i++) {
    double ai[3] = {0.0, 0.0, 0.0}, adi[3] = {0.0, 0.0, 0.0}, pot = 0.0;
    for (int j=0; j<N; j++) {
      if (i == j) continue;
      double dr[3], dv[3];
      for (int k=0; k<3; k++) { dr[k] = r[3*j+k] - r[3*i+k]; dv[k] = v[3*j+k] - v[3*i+k]; }
      const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
      const double invdrmag = 1.0/sqrt(drsqd);
      const double drdt = (dv[0]*dr[0] + dv[1]*dr[1] + dv[2]*dr[2])*invdrmag;
      const double inv3 = pow(invdrmag, 3);
      pot += m[j]*invdrmag;
      for (int k=0; k<3; k++) ai[k] += m[j]*dr[k]*inv3;
      for (int k=0; k<3; k++) adi[k] += m[j]*inv3*(dv[k] - 3.0*drdt*invdrmag*dr[k]);
    }
    for (int k=0; k<3; k++) { a[3*i+k] = ai[k]; adot[3*i+k] = adi[k]; }
    gpot[i] = pot;
  }
}

/* NbodyLeapfrogKDK::CalculateDirectSmoothedGravForces (src/Nbody/NbodyLeapfrogKDK.cpp:78-150) */
void orc_nbody_direct_smoothed(const orc_params* p, int N, const double* r, const double* v,
                               const double* m, const double* h, double* a, double* adot,
                               double* gpot) {
  okern kern; kern_init(&kern, p);
  #pragma omp parallel for schedule(dynamic, 16)
  for (int i=0; i<N; i++) {
    double ai[3] = {0.0, 0.0, 0.0}, adi[3] = {0.0, 0.0, 0.0}, pot = 0.0;
    for (int j=0; j<N; j++) {
      if (i == j) continue;
      double dr[3], dv[3];
      for (int k=0; k<3; k++) { dr[k] = r[3*j+k] - r[3*i+k]; dv[k] = v[3*j+k] - v[3*i+k]; }
      const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
      const double drmag = sqrt(drsqd) + SMALL_NUMBER;
      const double invdrmag = 1.0/drmag;
      const double invhmean = 2.0/(h[i] + h[j]);
      const double drdt = (dv[0]*dr[0] + dv[1]*dr[1] + dv[2]*dr[2])*invdrmag;
      const double paux = m[j]*invhmean*invhmean*K_wgrav(&kern, drmag*invhmean)*invdrmag;
      const double wmean = K_w0(&kern, drmag*invhmean)*powf(invhmean, 3);
      pot += m[j]*invhmean*K_wpot(&kern, drmag*invhmean);
      for (int k=0; k<3; k++) ai[k] += paux*dr[k];
      for (int k=0; k<3; k++)
        adi[k] += paux*dv[k] - 3.0*paux*drdt*invdrmag*dr[k] +
                  2.0*ORC_TWOPI*m[j]*drdt*wmean*invdrmag*dr[k];
    }
    for (int k=0; k<3; k++) { a[3*i+k] = ai[k]; adot[3*i+k] = adi[k]; }
    gpot[i] = pot;
  }
  kern_free(&kern);
}

/* ------------------------------------------------------------------------------------------ */
/* Device-tree twin: Morton keys, LBVH (Karras 2012), bottom-up moments, walk groups           */
/* ------------------------------------------------------------------------------------------ */
static inline uint64_t spread3(uint64_t x) {
  x &= 0x1fffffULL;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8))  & 0x100f00f00f00f00fULL;
  x = (x | (x << 4))  & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2))  & 0x1249249249249249ULL;
  return x;
}

static inline int delta_fn(const uint64_t* key, int N, int i, int64_t j) {
  if (j < 0 || j > N - 1) return -1;
  uint64_t a = key[i], b = key[j];
  if (a != b) return __builtin_clzll(a ^ b);
  return 64 + __builtin_clz((uint32_t) (i ^ (int) j));
}

int orc_tree_alloc(orc_tree* t, int N) {
  memset(t, 0, sizeof(*t));
  t->N = N;
  int nn = 2*N - 1;
  int ni = N - 1 > 0 ? N - 1 : 1;
  t->perm = malloc(sizeof(int32_t)*N);
  t->keys = malloc(sizeof(uint64_t)*N);
  t->left = malloc(sizeof(int32_t)*ni);
  t->right = malloc(sizeof(int32_t)*ni);
  t->first = malloc(sizeof(int32_t)*ni);
  t->last = malloc(sizeof(int32_t)*ni);
  t->parent = malloc(sizeof(int32_t)*nn);
  t->mass = malloc(sizeof(double)*nn);
  t->com = malloc(sizeof(double)*3*nn);
  t->quad = malloc(sizeof(double)*5*nn);
  t->bbmin = malloc(sizeof(double)*3*nn);
  t->bbmax = malloc(sizeof(double)*3*nn);
  t->hmax = malloc(sizeof(double)*nn);
  t->cdistsqd = malloc(sizeof(double)*nn);
  t->group_first = malloc(sizeof(int32_t)*N);
  t->group_count = malloc(sizeof(int32_t)*N);
  t->group_node = malloc(sizeof(int32_t)*N);
  return 0;
}
void orc_tree_free(orc_tree* t) {
  free(t->perm); free(t->keys); free(t->left); free(t->right); free(t->first); free(t->last);
  free(t->parent); free(t->mass); free(t->com); free(t->quad); free(t->bbmin); free(t->bbmax);
  free(t->hmax); free(t->cdistsqd); free(t->group_first); free(t->group_count);
  free(t->group_node);
  memset(t, 0, sizeof(*t));
}

typedef struct { uint64_t key; int32_t idx; } kv;
static int cmp_kv(const void* a, const void* b) {
  const kv* x = (const kv*) a; const kv* y = (const kv*) b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return (x->idx > y->idx) - (x->idx < y->idx);    /* stable: input order breaks ties */
}

static inline int node_count(const orc_tree* t, int n) {
  return n >= t->N - 1 ? 1 : t->last[n] - t->first[n] + 1;
}

/* merge children into internal node n, operation order as KDTree.cpp:950-1051 */
static void merge_node(const orc_params* p, orc_tree* t, int n) {
  const int L = t->left[n], R = t->right[n];
  const int ch[2] = {L, R};
  double m = t->mass[L] + t->mass[R];
  t->mass[n] = m;
  for (int k=0; k<3; k++) {
    t->com[3*n+k] = m > 0 ? (t->mass[L]*t->com[3*L+k] + t->mass[R]*t->com[3*R+k])/m : 0.0;
    t->bbmin[3*n+k] = fmin(t->bbmin[3*L+k], t->bbmin[3*R+k]);
    t->bbmax[3*n+k] = fmax(t->bbmax[3*L+k], t->bbmax[3*R+k]);
  }
  t->hmax[n] = fmax(t->hmax[L], t->hmax[R]);
  double q[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int c=0; c<2; c++) {
    const int C = ch[c];
    const double mi = t->mass[C];
    if (!(mi > 0)) continue;
    double dr[3];
    for (int k=0; k<3; k++) dr[k] = t->com[3*C+k] - t->com[3*n+k];
    const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
    for (int k=0; k<5; k++) q[k] += t->quad[5*C+k];
    q[0] += mi*(3.0*dr[0]*dr[0] - drsqd);
    q[1] += mi*3.0*dr[0]*dr[1];
    q[2] += mi*(3.0*dr[1]*dr[1] - drsqd);
    q[3] += mi*3.0*dr[2]*dr[0];
    q[4] += mi*3.0*dr[2]*dr[1];
  }
  for (int k=0; k<5; k++) t->quad[5*n+k] = q[k];
  double half[3];
  for (int k=0; k<3; k++) half[k] = 0.5*(t->bbmax[3*n+k] - t->bbmin[3*n+k]);
  const double rmaxsqd = half[0]*half[0] + half[1]*half[1] + half[2]*half[2];
  t->cdistsqd[n] = fmax(rmaxsqd, t->hmax[n]*t->hmax[n])/p->thetamaxsqd;
}

static void leaf_node(const orc_params* p, orc_tree* t, int l, const double* r,
                      const double* m, const double* h) {
  const int n = t->N - 1 + l;
  const int i = t->perm[l];
  t->mass[n] = m[i];
  for (int k=0; k<3; k++) {
    t->com[3*n+k] = r[3*i+k];
    t->bbmin[3*n+k] = r[3*i+k];
    t->bbmax[3*n+k] = r[3*i+k];
  }
  for (int k=0; k<5; k++) t->quad[5*n+k] = 0.0;
  t->hmax[n] = h[i];
  t->cdistsqd[n] = (h[i]*h[i])/p->thetamaxsqd;
}

/* post-order over internal nodes (children first) */
static void refit_all(const orc_params* p, orc_tree* t, int hmax_only) {
  const int N = t->N;
  if (N < 2) return;
  int* stack = malloc(sizeof(int)*2*N);
  char* seen = calloc(N - 1, 1);
  int sp = 0;
  stack[sp++] = 0;
  while (sp) {
    int n = stack[sp-1];
    if (!seen[n]) {
      seen[n] = 1;
      if (t->right[n] < N - 1) stack[sp++] = t->right[n];
      if (t->left[n] < N - 1) stack[sp++] = t->left[n];
    } else {
      sp--;
      if (hmax_only) {
        t->hmax[n] = fmax(t->hmax[t->left[n]], t->hmax[t->right[n]]);
        double half[3];
        for (int k=0; k<3; k++) half[k] = 0.5*(t->bbmax[3*n+k] - t->bbmin[3*n+k]);
        const double rmaxsqd = half[0]*half[0] + half[1]*half[1] + half[2]*half[2];
        t->cdistsqd[n] = fmax(rmaxsqd, t->hmax[n]*t->hmax[n])/p->thetamaxsqd;
      } else {
        merge_node(p, t, n);
      }
    }
  }
  free(stack); free(seen);
}

int orc_tree_build(const orc_params* p, orc_tree* t, const double* r, const double* m,
                   const double* h, int group_max) {
  const int N = t->N;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i=0; i<N; i++)
    for (int k=0; k<3; k++) {
      if (r[3*i+k] < lo[k]) lo[k] = r[3*i+k];
      if (r[3*i+k] > hi[k]) hi[k] = r[3*i+k];
    }
  double span = hi[0] - lo[0];
  if (hi[1] - lo[1] > span) span = hi[1] - lo[1];
  if (hi[2] - lo[2] > span) span = hi[2] - lo[2];
  if (!(span > 0.0)) span = 1.0;
  const double scale = 2097152.0/span;
  kv* a = malloc(sizeof(kv)*N);
  for (int i=0; i<N; i++) {
    uint64_t q[3];
    for (int k=0; k<3; k++) {
      double f = (r[3*i+k] - lo[k])*scale;
      q[k] = f >= 2097151.0 ? 2097151ULL : (uint64_t) f;
    }
    a[i].key = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
    a[i].idx = i;
  }
  qsort(a, N, sizeof(kv), cmp_kv);
  for (int l=0; l<N; l++) { t->perm[l] = a[l].idx; t->keys[l] = a[l].key; }
  free(a);

  /* internal nodes (Karras 2012, Fig. 4) */
  for (int i=0; i<N-1; i++) {
    const int d = (delta_fn(t->keys, N, i, (int64_t) i + 1) - delta_fn(t->keys, N, i, (int64_t) i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta_fn(t->keys, N, i, (int64_t) i - d);
    int64_t lmax = 2;
    while (delta_fn(t->keys, N, i, (int64_t) i + lmax*d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t s = lmax/2; s >= 1; s /= 2)
      if (delta_fn(t->keys, N, i, (int64_t) i + (l + s)*d) > dmin) l += s;
    const int64_t j = (int64_t) i + l*d;
    const int dnode = delta_fn(t->keys, N, i, j);
    int64_t s = 0, div = 2, step;
    do {
      step = (l + div - 1)/div;
      if (delta_fn(t->keys, N, i, (int64_t) i + (s + step)*d) > dnode) s += step;
      div *= 2;
    } while (step > 1);
    const int gamma = (int) ((int64_t) i + s*d + (d < 0 ? d : 0));
    const int first = (int) (i < j ? i : j), last = (int) (i < j ? j : i);
    t->first[i] = first; t->last[i] = last;
    t->left[i] = (first == gamma) ? (N - 1 + gamma) : gamma;
    t->right[i] = (last == gamma + 1) ? (N - 1 + gamma + 1) : gamma + 1;
  }
  t->parent[0] = -1;
  if (N == 1) t->parent[0] = -1;
  for (int i=0; i<N-1; i++) { t->parent[t->left[i]] = i; t->parent[t->right[i]] = i; }
  for (int l=0; l<N; l++) leaf_node(p, t, l, r, m, h);
  refit_all(p, t, 0);

  /* walk groups: maximal nodes with <= group_max particles, ordered by first particle */
  int ng = 0;
  const int root = N > 1 ? 0 : 0;
  int* stack = malloc(sizeof(int)*(2*N));
  int sp = 0;
  stack[sp++] = root;
  /* pre-order left-to-right visit yields groups in ascending first-particle order */
  while (sp) {
    int n = stack[--sp];
    int cnt = (N == 1) ? 1 : node_count(t, n);
    if (cnt <= group_max || n >= N - 1) {
      t->group_node[ng] = (N == 1) ? 0 : n;
      t->group_first[ng] = (N == 1) ? 0 : (n >= N - 1 ? n - (N - 1) : t->first[n]);
      t->group_count[ng] = cnt;
      ng++;
    } else {
      stack[sp++] = t->right[n];
      stack[sp++] = t->left[n];
    }
  }
  free(stack);
  t->ngroups = ng;
  return ng;
}

void orc_tree_refit_h(const orc_params* p, orc_tree* t, const double* h) {
  const int N = t->N;
  for (int l=0; l<N; l++) {
    const int n = N - 1 + l;
    const double hh = h[t->perm[l]];
    t->hmax[n] = hh;
    t->cdistsqd[n] = (hh*hh)/p->thetamaxsqd;
  }
  refit_all(p, t, 1);
}

/* leaf-node unified index of a sorted slot */
static inline int unified_count(const orc_tree* t, int n) { return node_count(t, n); }

/* Quadrupole cell force (NeighbourSearch.h:423-449), dr = rp - cell.r */
static inline void quad_force(const double* com, const double* q, double mc,
                              const double* rp, double* agrav, double* gpot) {
  double dr[3];
  for (int k=0; k<3; k++) dr[k] = rp[k] - com[k];
  const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER;
  const double invdrsqd = 1.0/drsqd;
  const double invdrmag = sqrt(invdrsqd);
  const double invdr5 = invdrsqd*invdrsqd*invdrmag;
  for (int k=0; k<3; k++) agrav[k] -= mc*dr[k]*invdrsqd*invdrmag;
  const double qscalar = q[0]*dr[0]*dr[0] + q[2]*dr[1]*dr[1] - (q[0] + q[2])*dr[2]*dr[2] +
    2.0*(q[1]*dr[0]*dr[1] + q[3]*dr[0]*dr[2] + q[4]*dr[1]*dr[2]);
  const double qfactor = 2.5*qscalar*invdr5*invdrsqd;
  agrav[0] += (q[0]*dr[0] + q[1]*dr[1] + q[3]*dr[2])*invdr5 - qfactor*dr[0];
  agrav[1] += (q[1]*dr[0] + q[2]*dr[1] + q[4]*dr[2])*invdr5 - qfactor*dr[1];
  agrav[2] += (q[3]*dr[0] + q[4]*dr[1] - (q[0] + q[2])*dr[2])*invdr5 - qfactor*dr[2];
  *gpot += mc*invdrmag + 0.5*qscalar*invdr5;
}

/* the cell term above on its own, for its known-answer tests (a point target outside a cell) */
void orc_cell_quad_force(const double* com, const double* q, double mc, const double* rp,
                         double* agrav, double* gpot) {
  quad_force(com, q, mc, rp, agrav, gpot);
}

/* Monopole cell force (ComputeCellMonopoleForces, NeighbourSearch.h:350-377): the cell's mass at
 * its centre of mass, dr = cell.r - rp */
static inline void mono_force(const double* com, double mc, const double* rp, double* agrav,
                              double* gpot) {
  double dr[3];
  for (int k=0; k<3; k++) dr[k] = com[k] - rp[k];
  const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2] + SMALL_NUMBER;
  const double invdrsqd = 1.0/drsqd;
  const double invdrmag = sqrt(invdrsqd);
  const double invdr3 = invdrsqd*invdrmag;
  *gpot += mc*invdrmag;
  for (int k=0; k<3; k++) agrav[k] += mc*dr[k]*invdr3;
}

void orc_cell_mono_force(const double* com, double mc, const double* rp, double* agrav,
                         double* gpot) {
  mono_force(com, mc, rp, agrav, gpot);
}

/* FastMultipoleForces (NeighbourSearch.h:478-760) with its second-order terms off (HoT = false):
 * the accepted cells' potential, acceleration and tidal tensor at one expansion point rc,
 * summed once per walk group and applied to each target as a first-order Taylor series about rc.
 * f[0] = pot, f[1..3] = acceleration (the reference's dphi takes the same increments as ac, so
 * one vector serves both), f[4..9] = tensor (xx, xy, yy, xz, yz, zz).
 * dr = cell.r - rc; quad = 0 adds the monopole part only (AddMonopoleContribution, :564-597),
 * quad = 1 the monopole and quadrupole parts (AddQuadrupoleContribution, :675-723). */
static void fast_add(double* f, const double* dr, double mc, const double* q, int quad) {
  {
    const double invdrmag = sqrt(1.0/(dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2]));
    const double invdrsqd = invdrmag*invdrmag;
    const double invdr3 = invdrsqd*invdrmag;
    f[0] += mc*invdrmag;
    const double mc3 = mc*invdr3;
    for (int k=0; k<3; k++) f[1+k] += mc3*dr[k];
    f[4] += mc3*(3.0*dr[0]*dr[0]*invdrsqd - 1);
    f[5] += mc3*(3.0*dr[0]*dr[1]*invdrsqd);
    f[6] += mc3*(3.0*dr[1]*dr[1]*invdrsqd - 1);
    f[7] += mc3*(3.0*dr[2]*dr[0]*invdrsqd);
    f[8] += mc3*(3.0*dr[2]*dr[1]*invdrsqd);
    f[9] += mc3*(3.0*dr[2]*dr[2]*invdrsqd - 1);
  }
  if (!quad) return;
  double d[3];
  for (int k=0; k<3; k++) d[k] = -dr[k];          /* rc - cell.r */
  const double drsqd = d[0]*d[0] + d[1]*d[1] + d[2]*d[2] + SMALL_NUMBER;
  const double invdrsqd = 1.0/drsqd;
  const double invdrmag = sqrt(invdrsqd);
  const double invdr5 = invdrsqd*invdrsqd*invdrmag;
  const double qzz = -(q[0] + q[2]);
  const double qscalar = q[0]*d[0]*d[0] + q[2]*d[1]*d[1] + qzz*d[2]*d[2] +
    2.0*(q[1]*d[0]*d[1] + q[3]*d[0]*d[2] + q[4]*d[1]*d[2]);
  const double qfactor = 2.5*qscalar*invdr5*invdrsqd;
  double qx[3] = {(q[0]*d[0] + q[1]*d[1] + q[3]*d[2])*invdr5,
                  (q[1]*d[0] + q[2]*d[1] + q[4]*d[2])*invdr5,
                  (q[3]*d[0] + q[4]*d[1] + qzz*d[2])*invdr5};
  f[0] += 0.5*qscalar*invdr5;
  for (int k=0; k<3; k++) f[1+k] += qx[k] - qfactor*d[k];
  for (int k=0; k<3; k++) qx[k] *= 5*invdrsqd;
  f[4] += qfactor*(7*d[0]*d[0]*invdrsqd - 1);
  f[5] += qfactor*(7*d[0]*d[1]*invdrsqd);
  f[6] += qfactor*(7*d[1]*d[1]*invdrsqd - 1);
  f[7] += qfactor*(7*d[0]*d[2]*invdrsqd);
  f[8] += qfactor*(7*d[1]*d[2]*invdrsqd);
  f[9] += qfactor*(7*d[2]*d[2]*invdrsqd - 1);
  f[4] -= 2.0*qx[0]*d[0] - q[0]*invdr5;
  f[5] -= qx[0]*d[1] + qx[1]*d[0] - q[1]*invdr5;
  f[6] -= 2.0*qx[1]*d[1] - q[2]*invdr5;
  f[7] -= qx[0]*d[2] + qx[2]*d[0] - q[3]*invdr5;
  f[8] -= qx[1]*d[2] + qx[2]*d[1] - q[4]*invdr5;
  f[9] -= 2.0*qx[2]*d[2] - qzz*invdr5;
}

/* ApplyForcesTaylor (NeighbourSearch.h:735-760, first order): d = r - rc */
static inline void fast_apply(const double* f, const double* d, double* agrav, double* gpot) {
  agrav[0] += f[1] + f[4]*d[0] + f[5]*d[1] + f[7]*d[2];
  agrav[1] += f[2] + f[5]*d[0] + f[6]*d[1] + f[8]*d[2];
  agrav[2] += f[3] + f[7]*d[0] + f[8]*d[1] + f[9]*d[2];
  *gpot += f[0] + f[1]*d[0] + f[2]*d[1] + f[3]*d[2];
}

/* the fast expansion of n cells (com n x 3 relative to rc, q n x 5, m n) evaluated at the
 * points d (np x 3, relative to rc): for its known-answer tests */
void orc_fast_multipole(int n, const double* dcom, const double* q, const double* m, int quad,
                        int np, const double* d, double* coef, double* agrav, double* gpot) {
  for (int k=0; k<10; k++) coef[k] = 0.0;
  for (int c=0; c<n; c++) fast_add(coef, &dcom[3*c], m[c], &q[5*c], quad);
  for (int i=0; i<np; i++) {
    agrav[3*i] = agrav[3*i+1] = agrav[3*i+2] = 0.0;
    gpot[i] = 0.0;
    fast_apply(coef, &d[3*i], &agrav[3*i], &gpot[i]);
  }
}

#define GRAV_NEIB_SLACK (1.0 + 1.0e-4)
#define TWOTHIRDS 0.66666666666666666666666

/* The walking group's factor of the gadget2 and eigenmac criteria, from its particles' state at
 * the tree stock (KDTree.cpp:900-903, the walking leaf cell's amin and macfactor): gadget2 the
 * smallest |atree|, eigenmac the largest gpot^(-2/3).  NULL for the geometric criterion. */
static double* group_mac_factors(const orc_params* p, const orc_tree* t, const double* atree,
                                 const double* gpot) {
  if (p->gravity_mac == 0) return NULL;
  double* out = malloc(sizeof(double)*(t->ngroups > 0 ? t->ngroups : 1));
  for (int g=0; g<t->ngroups; g++) {
    double v = p->gravity_mac == 1 ? 9.9e30 : 0.0;
    for (int tt=0; tt<t->group_count[g]; tt++) {
      const int i = t->perm[t->group_first[g] + tt];
      if (p->gravity_mac == 1) {
        const double* a = &atree[3*i];
        v = fmin(v, sqrt(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]));
      } else {
        v = fmax(v, pow(gpot[i], -TWOTHIRDS));
      }
    }
    out[g] = v;
  }
  return out;
}

/* Tree::open_cell_for_gravity (Tree.h:413-432) past its geometric test: gadget2 opens a cell
 * whose m rmax^2 / d^4 exceeds macerror |a|; eigenmac one nearer than its (0.5 lambda /
 * macerror)^(2/3) (KDTree.cpp:1058-1079, lambda bounding the quadrupole's largest eigenvalue)
 * times the group's factor.  d^2 is the group box's nearest-point distance, as in the geometric
 * test. */
static int mac_opens(const orc_params* p, const double* half, double mc, const double* q,
                     double dmac2, double gfac) {
  if (p->gravity_mac == 1) {
    const double rmax = sqrt(half[0]*half[0] + half[1]*half[1] + half[2]*half[2]);
    return dmac2*dmac2*gfac*p->macerror < rmax*rmax*mc;
  }
  if (p->gravity_mac == 2) {
    const double pq = q[0]*q[2] - (q[0] + q[2])*(q[0] + q[2]) - q[1]*q[1] - q[3]*q[3] -
                      q[4]*q[4];
    const double lambda = pq >= 0.0 ? 0.0 : 2.0*sqrt(-pq/3.0);
    const double mac = pow(0.5*lambda/p->macerror, 0.66666666666666);
    return dmac2 < mac*gfac;
  }
  return 0;
}

/* the criterion above for one cell (box half-widths half, mass mc, moments q[5]), for its
 * known-answer tests */
int orc_cell_mac_opens(const orc_params* p, const double* half, double mc, const double* q,
                       double dmac2, double gfac) {
  return mac_opens(p, half, mc, q, dmac2, gfac);
}

/* act (optional, input order): groups without an active particle are skipped (counts 0) and
 * only active targets are written */
static void tree_gravity_masked(const orc_params* p, const orc_tree* t, const double* r,
                                const double* m, const double* h, const double* zeta,
                                const double* hfactor, double* atree, double* gpot,
                                int32_t* ncell, int32_t* nnear, const unsigned char* act,
                                const double* gmacv) {
  okern kern; kern_init(&kern, p);
  obox b; box_init(&b, p, 0);
  const int N = t->N;
  const double kr = kern.kernrange, krsqd = kern.kernrangesqd;
  const double wpot0 = K_wpot(&kern, 0.0);
  const int nleaf = p->nleafmax;
  #pragma omp parallel
  {
    int* stack = malloc(sizeof(int)*(2*N + 2));
    int* cells = malloc(sizeof(int)*(2*N + 2));
    int* near = malloc(sizeof(int)*(N + 2));
    #pragma omp for schedule(dynamic, 4)
    for (int g=0; g<t->ngroups; g++) {
      if (act) {
        int any = 0;
        for (int tt=0; tt<t->group_count[g]; tt++) any |= act[t->perm[t->group_first[g] + tt]];
        if (!any) {
          if (ncell) ncell[g] = 0;
          if (nnear) nnear[g] = 0;
          continue;
        }
      }
      const int G = t->group_node[g];
      double cG[3], halfG[3];
      for (int k=0; k<3; k++) {
        cG[k] = 0.5*(t->bbmin[3*G+k] + t->bbmax[3*G+k]);
        halfG[k] = 0.5*(t->bbmax[3*G+k] - t->bbmin[3*G+k]);
      }
      const double hmaxG = t->hmax[G];
      /* fast multipoles: the expansion point is the group's centre of mass (as the reference's
         is its walking leaf cell's, GradhSphTree.cpp:587-593); an accepted cell must also lie
         beyond the group's own extent about that point divided by theta, so that the Taylor
         series converges for every target of the group */
      const int fast = p->multipole >= 2;
      double rcr[3], rg2 = 0.0;
      for (int k=0; k<3; k++) {
        rcr[k] = t->com[3*G+k] - cG[k];
        const double e = fabs(rcr[k]) + halfG[k];
        rg2 += e*e;
      }
      const double rg2t = rg2/p->thetamaxsqd;
      const double gfac = gmacv ? gmacv[g] : 0.0;
      int sp = 0, nc = 0, np = 0;
      stack[sp++] = 0;   /* root: internal 0, or the single leaf when N == 1 */
      while (sp) {
        const int c = stack[--sp];
        const int cnt = (N == 1) ? 1 : unified_count(t, c);
        double dr[3], half[3];
        for (int k=0; k<3; k++) {
          dr[k] = 0.5*(t->bbmin[3*c+k] + t->bbmax[3*c+k]) - cG[k];
          half[k] = 0.5*(t->bbmax[3*c+k] - t->bbmin[3*c+k]);
        }
        nearest(&b, dr);
        double dbox2 = 0.0;
        for (int k=0; k<3; k++) {
          double gk = fabs(dr[k]) - half[k] - halfG[k];
          if (gk > 0.0) dbox2 += gk*gk;
        }
        const double hr = kr*fmax(hmaxG, t->hmax[c]);
        int open = 0, leafadd = 0;
        /* relative slack on the kernel-range box test: keeps every scatter-gather pair out of
           accepted multipoles whatever the rounding of the per-pair test */
        if (dbox2 < hr*hr*GRAV_NEIB_SLACK) {
          if (cnt <= nleaf || c >= N - 1 || N == 1) leafadd = 1; else open = 1;
        } else {
          double dmac2 = 0.0;
          for (int k=0; k<3; k++) {
            double ek = fabs(dr[k]) - halfG[k];
            if (ek > 0.0) dmac2 += ek*ek;
          }
          int accept = dmac2 >= t->cdistsqd[c];
          if (accept && gmacv)
            accept = !mac_opens(p, half, t->mass[c], &t->quad[5*c], dmac2, gfac);
          if (accept && fast) {
            double dc[3];
            for (int k=0; k<3; k++) dc[k] = t->com[3*c+k] - cG[k];
            nearest(&b, dc);
            double dcom2 = 0.0;
            for (int k=0; k<3; k++) { dc[k] -= rcr[k]; dcom2 += dc[k]*dc[k]; }
            accept = dcom2 >= rg2t;
          }
          if (accept) {
            if (cnt == 1) leafadd = 1; else cells[nc++] = c;
          } else {
            if (cnt <= nleaf || c >= N - 1 || N == 1) leafadd = 1; else open = 1;
          }
        }
        if (leafadd) {
          if (N == 1 || c >= N - 1) near[np++] = (N == 1) ? 0 : c - (N - 1);
          else for (int s=t->first[c]; s<=t->last[c]; s++) near[np++] = s;
        } else if (open) {
          stack[sp++] = t->right[c];
          stack[sp++] = t->left[c];
        }
      }
      if (ncell) ncell[g] = nc;
      if (nnear) nnear[g] = np;
      double fc[10] = {0.0};
      if (fast) {
        for (int q=0; q<nc; q++) {
          const int c = cells[q];
          double dc[3];
          for (int k=0; k<3; k++) dc[k] = t->com[3*c+k] - cG[k];
          nearest(&b, dc);
          for (int k=0; k<3; k++) dc[k] -= rcr[k];
          fast_add(fc, dc, t->mass[c], &t->quad[5*c], p->multipole == 3);
        }
      }
      for (int tt=0; tt<t->group_count[g]; tt++) {
        const int i = t->perm[t->group_first[g] + tt];
        if (act && !act[i]) continue;
        double acc[3] = {0.0, 0.0, 0.0};
        double pot = (m[i]/h[i])*wpot0;
        for (int q=0; q<np; q++) {
          const int j = t->perm[near[q]];
          double d[3] = {r[3*j] - r[3*i], r[3*j+1] - r[3*i+1], r[3*j+2] - r[3*i+2]};
          nearest(&b, d);
          grav_pair(&kern, krsqd, d, m[j], h[i], h[j], zeta[i], zeta[j], hfactor[i],
                    hfactor[j], acc, &pot);
        }
        for (int q=0; q<nc && !fast; q++) {
          const int c = cells[q];
          double rp[3] = {r[3*i], r[3*i+1], r[3*i+2]};
          if (b.periodic) {
            double d[3];
            for (int k=0; k<3; k++) d[k] = t->com[3*c+k] - rp[k];
            nearest(&b, d);
            for (int k=0; k<3; k++) rp[k] = t->com[3*c+k] - d[k];
          }
          if (p->multipole == 1) mono_force(&t->com[3*c], t->mass[c], rp, acc, &pot);
          else quad_force(&t->com[3*c], &t->quad[5*c], t->mass[c], rp, acc, &pot);
        }
        if (fast) {
          const double d[3] = {r[3*i] - t->com[3*G], r[3*i+1] - t->com[3*G+1],
                               r[3*i+2] - t->com[3*G+2]};
          fast_apply(fc, d, acc, &pot);
        }
        for (int k=0; k<3; k++) atree[3*i+k] = acc[k];
        gpot[i] = pot;
      }
    }
    free(stack); free(cells); free(near);
  }
  kern_free(&kern);
}

void orc_tree_gravity(const orc_params* p, const orc_tree* t, const double* r,
                      const double* m, const double* h, const double* zeta,
                      const double* hfactor, double* atree, double* gpot,
                      int32_t* ncell, int32_t* nnear, const double* atree_prev,
                      const double* gpot_prev) {
  double* gm = group_mac_factors(p, t, atree_prev, gpot_prev);
  tree_gravity_masked(p, t, r, m, h, zeta, hfactor, atree, gpot, ncell, nnear, NULL, gm);
  free(gm);
}

/* ------------------------------------------------------------------------------------------ */
/* Integrator, timesteps, diagnostics                                                          */
/* ------------------------------------------------------------------------------------------ */
void orc_advance(int N, double dt, int energy_integration, const double* r0,
                 const double* v0, const double* a0, const double* u0, const double* dudt0,
                 double* r, double* v, double* u) {
  #pragma omp parallel for
  for (int i=0; i<N; i++) {
    for (int k=0; k<3; k++) r[3*i+k] = r0[3*i+k] + v0[3*i+k]*dt + 0.5*a0[3*i+k]*dt*dt;
    for (int k=0; k<3; k++) v[3*i+k] = v0[3*i+k] + a0[3*i+k]*dt;
    if (energy_integration) u[i] = u0[i] + dudt0[i]*dt;
  }
}

void orc_end_step(int N, double dtstep, int energy_integration, const double* a,
                  const double* dudt, double* v, double* u, double* r0, double* v0,
                  double* a0, double* u0, double* dudt0, const double* r) {
  #pragma omp parallel for
  for (int i=0; i<N; i++) {
    for (int k=0; k<3; k++) v[3*i+k] += 0.5*dtstep*(a[3*i+k] - a0[3*i+k]);
    for (int k=0; k<3; k++) r0[3*i+k] = r[3*i+k];
    for (int k=0; k<3; k++) v0[3*i+k] = v[3*i+k];
    for (int k=0; k<3; k++) a0[3*i+k] = a[3*i+k];
    if (energy_integration) {
      u[i] += 0.5*(dudt[i] - dudt0[i])*dtstep;
      if (u[i] <= 0.0) u[i] = u0[i] + dudt0[i]*dtstep;
      u0[i] = u[i];
      dudt0[i] = dudt[i];
    }
  }
}

double orc_timesteps(const orc_params* p, int N, const double* h, const double* sound,
                     const double* div_v, const double* a, const double* u,
                     const double* dudt, double* dt_out) {
  double dtmin = 9.9e50;
  #pragma omp parallel for reduction(min:dtmin)
  for (int i=0; i<N; i++) {
    double timestep;
    if (p->hydro_forces)
      timestep = p->courant_mult*h[i]/(sound[i] + h[i]*fabs(div_v[i]) + SMALL_NUMBER_DP);
    else
      timestep = p->courant_mult*h[i]/(h[i]*fabs(div_v[i]) + SMALL_NUMBER_DP);
    double amag = sqrt(a[3*i]*a[3*i] + a[3*i+1]*a[3*i+1] + a[3*i+2]*a[3*i+2]);
    timestep = fmin(timestep, p->accel_mult*sqrt(h[i]/(amag + SMALL_NUMBER_DP)));
    if (p->energy_integration)
      timestep = fmin(timestep, p->energy_mult*(double) (u[i]/(fabs(dudt[i]) + SMALL_NUMBER)));
    if (dt_out) dt_out[i] = timestep;
    if (timestep < dtmin) dtmin = timestep;
  }
  return dtmin;
}

void orc_wrap_periodic(const orc_params* p, int N, double* r) {
  if (!p->periodic) return;
  for (int i=0; i<N; i++)
    for (int k=0; k<3; k++) {
      double size = p->boxmax[k] - p->boxmin[k];
      if (r[3*i+k] < p->boxmin[k]) r[3*i+k] += size;
      if (r[3*i+k] > p->boxmax[k]) r[3*i+k] -= size;
    }
}

void orc_diagnostics(const orc_params* p, int N, const double* r, const double* v,
                     const double* m, const double* u, const double* gpot, double* out) {
  double mtot = 0, ketot = 0, utot = 0, gpetot = 0, mom[3] = {0, 0, 0}, ang[3] = {0, 0, 0};
  for (int i=0; i<N; i++) {
    const double* ri = &r[3*i]; const double* vi = &v[3*i];
    mtot += m[i];
    ketot += m[i]*(vi[0]*vi[0] + vi[1]*vi[1] + vi[2]*vi[2]);
    utot += m[i]*u[i];
    gpetot -= m[i]*gpot[i];
    for (int k=0; k<3; k++) mom[k] += m[i]*vi[k];
  }
  for (int i=0; i<N; i++) {
    const double* ri = &r[3*i]; const double* vi = &v[3*i];
    ang[0] += m[i]*(ri[1]*vi[2] - ri[2]*vi[1]);
    ang[1] += m[i]*(ri[2]*vi[0] - ri[0]*vi[2]);
    ang[2] += m[i]*(ri[0]*vi[1] - ri[1]*vi[0]);
  }
  ketot *= 0.5;
  gpetot *= 0.5;
  double Etot = ketot;
  if (p->hydro_forces) Etot += utot;
  if (p->self_gravity) Etot += gpetot;
  out[0] = Etot; out[1] = ketot; out[2] = utot; out[3] = gpetot;
  for (int k=0; k<3; k++) { out[4+k] = mom[k]; out[7+k] = ang[k]; }
  out[10] = mtot;
}

/* ------------------------------------------------------------------------------------------ */
/* Full-step driver                                                                            */
/* ------------------------------------------------------------------------------------------ */
#define ORC_INVLOGETWO 1.44269504088896   /* Constants.h:81 */

struct orc_sim {
  orc_params p;
  int N;
  double t, timestep, tlast;
  double *r, *v, *a, *atree, *m, *h, *rho, *u, *dudt, *div_v, *gpot, *invomega, *zeta;
  double *sound, *pressure, *hfactor, *r0, *v0, *a0, *u0, *dudt0;
  /* atree and gpot at the step's tree stock, for the gadget2 and eigenmac criteria */
  double *mac_a, *mac_pot;
  /* block timesteps (Particle.h:140-167, Simulation.h:187-215) */
  int n, nresync, level_max, level_step;
  double dt_max;
  int *level, *levelneib, *nstep, *nlast;
  double *ptlast, *pdt, *dt_next, *dtw;
  unsigned char *active, *endstep;
  long long nfinished;   /* particles that ended a step in the last MainLoop */
  /* stars beside the gas (orc_sim_set_stars) */
  int ns, star_soft, star_cons;
  double nbody_mult;
  double *sr, *sv, *sa, *sadot, *sr0, *sv0, *sa0, *sm, *sh, *sgpot;
  orc_stars st;
};

static const orc_stars* sim_stars(orc_sim* s) {
  if (s->ns <= 0) return NULL;
  s->st.n = s->ns; s->st.r = s->sr; s->st.m = s->sm; s->st.h = s->sh;
  s->st.softening = s->star_soft; s->st.conservative = s->star_cons;
  return &s->st;
}

int orc_sim_set_stars(orc_sim* s, int n, const double* r, const double* v, const double* m,
                      const double* h, int softening, int conservative, double nbody_mult) {
  if (n < 0) return -1;
  double** f[] = {&s->sr, &s->sv, &s->sa, &s->sadot, &s->sr0, &s->sv0, &s->sa0};
  for (unsigned k=0; k<sizeof(f)/sizeof(f[0]); k++) {
    free(*f[k]);
    *f[k] = calloc((size_t) 3*(n > 0 ? n : 1), sizeof(double));
  }
  free(s->sm); free(s->sh); free(s->sgpot);
  s->sm = calloc((size_t) (n > 0 ? n : 1), sizeof(double));
  s->sh = calloc((size_t) (n > 0 ? n : 1), sizeof(double));
  s->sgpot = calloc((size_t) (n > 0 ? n : 1), sizeof(double));
  s->ns = n; s->star_soft = softening; s->star_cons = conservative; s->nbody_mult = nbody_mult;
  if (n > 0) {
    memcpy(s->sr, r, sizeof(double)*3*n); memcpy(s->sr0, r, sizeof(double)*3*n);
    memcpy(s->sv, v, sizeof(double)*3*n); memcpy(s->sv0, v, sizeof(double)*3*n);
    memcpy(s->sm, m, sizeof(double)*n); memcpy(s->sh, h, sizeof(double)*n);
  }
  return 0;
}

orc_sim* orc_sim_create(const orc_params* p, int N, const double* r, const double* v,
                        const double* m, const double* u) {
  orc_sim* s = calloc(1, sizeof(orc_sim));
  s->p = *p; s->N = N;
#define ALLOC(f, n) s->f = calloc((size_t) (n), sizeof(*s->f))
  ALLOC(r, 3*N); ALLOC(v, 3*N); ALLOC(a, 3*N); ALLOC(atree, 3*N); ALLOC(m, N); ALLOC(h, N);
  ALLOC(rho, N); ALLOC(u, N); ALLOC(dudt, N); ALLOC(div_v, N); ALLOC(gpot, N);
  ALLOC(invomega, N); ALLOC(zeta, N); ALLOC(sound, N); ALLOC(pressure, N); ALLOC(hfactor, N);
  ALLOC(r0, 3*N); ALLOC(v0, 3*N); ALLOC(a0, 3*N); ALLOC(u0, N); ALLOC(dudt0, N);
  ALLOC(level, N); ALLOC(levelneib, N); ALLOC(nstep, N); ALLOC(nlast, N);
  ALLOC(ptlast, N); ALLOC(pdt, N); ALLOC(dt_next, N); ALLOC(dtw, N);
  ALLOC(active, N); ALLOC(endstep, N); ALLOC(mac_a, 3*N); ALLOC(mac_pot, N);
#undef ALLOC
  memcpy(s->r, r, sizeof(double)*3*N);
  memcpy(s->v, v, sizeof(double)*3*N);
  memcpy(s->m, m, sizeof(double)*N);
  memcpy(s->u, u, sizeof(double)*N);
  return s;
}

void orc_sim_destroy(orc_sim* s) {
  if (!s) return;
  free(s->r); free(s->v); free(s->a); free(s->atree); free(s->m); free(s->h); free(s->rho);
  free(s->u); free(s->dudt); free(s->div_v); free(s->gpot); free(s->invomega); free(s->zeta);
  free(s->sound); free(s->pressure); free(s->hfactor); free(s->r0); free(s->v0); free(s->a0);
  free(s->u0); free(s->dudt0);
  free(s->level); free(s->levelneib); free(s->nstep); free(s->nlast); free(s->ptlast);
  free(s->pdt); free(s->dt_next); free(s->dtw); free(s->active); free(s->endstep);
  free(s->mac_a); free(s->mac_pot);
  free(s->sr); free(s->sv); free(s->sa); free(s->sadot); free(s->sr0); free(s->sv0);
  free(s->sa0); free(s->sm); free(s->sh); free(s->sgpot);
  free(s);
}

/* the tree stock's copy of atree and gpot (KDTree.cpp:900-903), taken where the step builds its
 * tree: before the setup's first force pass and at the start of every MainLoop */
static void mac_snapshot(orc_sim* s) {
  memcpy(s->mac_a, s->atree, sizeof(double)*3*s->N);
  memcpy(s->mac_pot, s->gpot, sizeof(double)*s->N);
}

/* Sph::ZeroAccelerations (Sph.cpp:126-140), active particles */
static void zero_active(orc_sim* s, const unsigned char* act) {
  for (int i=0; i<s->N; i++) {
    if (act && !act[i]) continue;
    s->levelneib[i] = 0;
    s->div_v[i] = 0.0; s->dudt[i] = 0.0; s->gpot[i] = 0.0;
    for (int k=0; k<3; k++) { s->a[3*i+k] = 0.0; s->atree[3*i+k] = 0.0; }
  }
}

/* density, thermal state and forces of the active particles (all when act is NULL):
 * the body of SphSimulation::MainLoop's force loop (SphSimulation.cpp:653-720) */
static int sim_forces(orc_sim* s, const unsigned char* act, int with_h) {
  const int N = s->N;
  const orc_params* p = &s->p;
  if (with_h && compute_h_masked(p, N, s->r, s->m, 0, s->h, s->rho, s->invomega, s->zeta,
                                 s->hfactor, s->u, s->sound, s->pressure, act, sim_stars(s)))
    return -1;
  zero_active(s, act);
  /* ComputeThermalProperties for every particle (nradstep = 1) */
  if (act) for (int i=0; i<N; i++) eos_thermal(p, s->rho[i], &s->u[i], &s->sound[i],
                                               &s->pressure[i]);
  hydro_forces_masked(p, p->self_gravity ? 1 : 0, N, s->r, s->v, s->m, s->h, s->rho,
                      s->invomega, s->pressure, s->sound, s->hfactor, s->a, s->dudt, s->div_v,
                      act, s->level, s->levelneib);
  if (p->self_gravity && p->grav_direct) {
    direct_gravity_masked(p, N, s->r, s->m, s->h, s->zeta, s->hfactor, s->atree, s->gpot, act);
    for (int i=0; i<N; i++) {
      if (act && !act[i]) continue;
      for (int k=0; k<3; k++) s->a[3*i+k] += s->atree[3*i+k];
    }
  } else if (p->self_gravity) {
    orc_tree tr;
    orc_tree_alloc(&tr, N);
    orc_tree_build(p, &tr, s->r, s->m, s->h, 64);
    double* gm = group_mac_factors(p, &tr, s->mac_a, s->mac_pot);
    tree_gravity_masked(p, &tr, s->r, s->m, s->h, s->zeta, s->hfactor, s->atree, s->gpot,
                        NULL, NULL, act, gm);
    free(gm);
    orc_tree_free(&tr);
    for (int i=0; i<N; i++) {
      if (act && !act[i]) continue;
      for (int k=0; k<3; k++) s->a[3*i+k] += s->atree[3*i+k];
    }
  }
  if (s->ns > 0) {
    /* the stars' pull on the gas, into a and atree (GradhSphTree.cpp:600-617) */
    double* ga = calloc((size_t) 3*N, sizeof(double));
    orc_gas_from_stars(p, N, s->r, s->h, sim_stars(s), ga, s->gpot);
    for (int i=0; i<3*N; i++) { s->a[i] += ga[i]; s->atree[i] += ga[i]; }
    free(ga);
  }
  return 0;
}

/* the N-body iteration of the step (SphSimulation.cpp:775-807): the stars' a, adot, gpot
 * cleared, the gas's pull (self-gravitating runs), then the star-star sums */
static void star_forces(orc_sim* s) {
  const int ns = s->ns;
  if (ns <= 0) return;
  if (s->p.self_gravity) {
    orc_stars_from_gas(&s->p, sim_stars(s), s->N, s->r, s->m, s->h, s->sa, s->sgpot);
  } else {
    memset(s->sa, 0, sizeof(double)*3*ns);
    memset(s->sgpot, 0, sizeof(double)*ns);
  }
  memset(s->sadot, 0, sizeof(double)*3*ns);
  double* a = calloc((size_t) 3*ns, sizeof(double));
  double* ad = calloc((size_t) 3*ns, sizeof(double));
  double* gp = calloc((size_t) ns, sizeof(double));
  if (s->star_soft) orc_nbody_direct_smoothed(&s->p, ns, s->sr, s->sv, s->sm, s->sh, a, ad, gp);
  else orc_nbody_direct(ns, s->sr, s->sv, s->sm, a, ad, gp);
  for (int i=0; i<3*ns; i++) { s->sa[i] += a[i]; s->sadot[i] += ad[i]; }
  for (int i=0; i<ns; i++) s->sgpot[i] += gp[i];
  free(a); free(ad); free(gp);
}

/* NbodyLeapfrogKDK::Timestep (NbodyLeapfrogKDK.cpp:386-401), minimum over the stars */
static double star_timestep(const orc_sim* s) {
  double dt = 1.0e300;
  for (int i=0; i<s->ns; i++) {
    const double* a = &s->sa[3*i];
    const double amag = sqrt(a[0]*a[0] + a[1]*a[1] + a[2]*a[2]);
    dt = fmin(dt, s->nbody_mult*sqrt(s->sh[i]/(amag + 1.0e-50)));
  }
  return dt;
}

/* NbodyLeapfrogKDK::EndTimestep (:333-370) */
static void star_end(orc_sim* s) {
  memcpy(s->sr0, s->sr, sizeof(double)*3*s->ns);
  memcpy(s->sv0, s->sv, sizeof(double)*3*s->ns);
  memcpy(s->sa0, s->sa, sizeof(double)*3*s->ns);
}

static int ipow2(int e) { return (int) pow(2.0, (double) e); }

/* InlineFuncs.h:550-557 */
static int timestep_level(double dt, double dt_max) {
  const int level = (int) (ORC_INVLOGETWO*log(dt_max/dt)) + 1;
  return level > 0 ? level : 0;
}

/* Simulation::ComputeBlockTimesteps (Simulation.cpp:1764-2197), hydro particles only, KDK
 * (integration_step = 1), no sinks, sph_single_timestep = 0 */
static void block_timesteps(orc_sim* s) {
  const int N = s->N;
  const orc_params* p = &s->p;
  const int ldm = p->level_diff_max;
  double dmin = orc_timesteps(p, N, s->h, s->sound, s->div_v, s->a, s->u, s->dudt, s->dtw);
  if (s->n == s->nresync) {
    s->n = 0;
    s->timestep = dmin;
    s->level_max = p->nlevels - 1;
    s->level_step = s->level_max;
    s->dt_max = s->timestep*pow(2.0, s->level_max);
    for (int i=0; i<N; i++) {
      int level = timestep_level(s->dtw[i], s->dt_max);
      if (level > s->level_max) level = s->level_max;
      s->level[i] = level;
      s->levelneib[i] = level;
      s->nstep[i] = ipow2(s->level_step - level);
      s->nlast[i] = s->n;
      s->dt_next[i] = s->nstep[i]*s->timestep;
      s->endstep[i] = 1;
    }
    s->nresync = ipow2(s->level_step);
    s->timestep = s->dt_max/(double) s->nresync;
    return;
  }
  const int level_max_old = s->level_max;
  int level_max = 0;
  for (int i=0; i<N; i++) {
    const int n = s->n;
    if (n - s->nlast[i] == s->nstep[i] && s->nstep[i] != ipow2(s->level_step - s->level[i])) {
      /* step ended early by CheckTimesteps (Saitoh & Makino limiter) */
      int level = timestep_level(s->dtw[i], s->dt_max);
      if (s->levelneib[i] - ldm > level) level = s->levelneib[i] - ldm;
      if (level > s->level[i]) s->level[i] = level;
      s->levelneib[i] = s->level[i];
      s->nlast[i] = n;
      s->nstep[i] = ipow2(s->level_step - s->level[i]);
      s->dt_next[i] = s->nstep[i]*s->timestep;
      s->endstep[i] = 1;
    } else if (n - s->nlast[i] == s->nstep[i]) {
      const int nstep = s->nstep[i], last_level = s->level[i];
      int level = timestep_level(s->dtw[i], s->dt_max);
      if (s->levelneib[i] - ldm > level) level = s->levelneib[i] - ldm;
      if (level < last_level && last_level > 1 && n%(2*nstep) == 0) s->level[i] = last_level - 1;
      else if (level > last_level) s->level[i] = level;
      else s->level[i] = last_level;
      s->levelneib[i] = level;
      s->nlast[i] = n;
      s->nstep[i] = ipow2(s->level_step - s->level[i]);
      s->dt_next[i] = s->nstep[i]*s->timestep;
      s->endstep[i] = 1;
    }
    if (s->level[i] > level_max) level_max = s->level[i];
  }
  const int istep = ipow2(s->level_step - level_max_old + 1);
  if (level_max > level_max_old) {
    const int nfactor = ipow2(level_max - level_max_old);
    s->n *= nfactor;
    for (int i=0; i<N; i++) { s->nstep[i] *= nfactor; s->nlast[i] *= nfactor; }
  } else if (level_max <= level_max_old - 1 && level_max_old > 1 && s->n%istep == 0) {
    level_max = level_max_old - 1;
    const int nfactor = ipow2(level_max_old - level_max);
    s->n /= nfactor;
    for (int i=0; i<N; i++) { s->nlast[i] /= nfactor; s->nstep[i] /= nfactor; }
  } else {
    level_max = level_max_old;
  }
  s->level_max = level_max;
  s->level_step = level_max;
  s->nresync = ipow2(s->level_step);
  s->timestep = s->dt_max/(double) s->nresync;
  for (int i=0; i<N; i++)
    if (s->nlast[i] == s->n) s->nstep[i] = ipow2(s->level_step - s->level[i]);
}

/* SphLeapfrogKDK::CheckTimesteps (SphLeapfrogKDK.cpp:284-329) */
static int check_timesteps(orc_sim* s) {
  int activecount = 0;
  for (int i=0; i<s->N; i++) {
    const int dn = s->n - s->nlast[i];
    if (dn == s->nstep[i]) continue;
    if (s->levelneib[i] - s->level[i] > s->p.level_diff_max) {
      const int level_new = s->levelneib[i] - s->p.level_diff_max;
      const int nnewstep = ipow2(s->level_step - level_new);
      if (nnewstep > 0 && dn%nnewstep == 0) {
        if (dn > 0) s->nstep[i] = dn;
        s->level[i] = level_new;
        s->active[i] = 1;
        activecount++;
      }
    }
  }
  return activecount;
}

/* SphLeapfrogKDK::EndTimestep (SphLeapfrogKDK.cpp:219-272), particles flagged end_timestep */
static void end_timestep(orc_sim* s) {
  const int ei = s->p.energy_integration;
  long long nf = 0;
  for (int i=0; i<s->N; i++) {
    if (!s->endstep[i]) continue;
    nf++;
    const double dt = s->pdt[i];
    for (int k=0; k<3; k++) s->v[3*i+k] += 0.5*dt*(s->a[3*i+k] - s->a0[3*i+k]);
    for (int k=0; k<3; k++) s->r0[3*i+k] = s->r[3*i+k];
    for (int k=0; k<3; k++) s->v0[3*i+k] = s->v[3*i+k];
    for (int k=0; k<3; k++) s->a0[3*i+k] = s->a[3*i+k];
    if (ei) {
      s->u[i] += 0.5*(s->dudt[i] - s->dudt0[i])*dt;
      if (s->u[i] <= 0.0) s->u[i] = s->u0[i] + s->dudt0[i]*dt;
      s->u0[i] = s->u[i];
      s->dudt0[i] = s->dudt[i];
    }
    s->nlast[i] = s->n;
    s->ptlast[i] = s->t;
    s->pdt[i] = s->dt_next[i];
    s->dt_next[i] = 0.0;
    s->active[i] = 0;
    s->endstep[i] = 0;
  }
  s->nfinished = nf;
}

/* SphSimulation.cpp:204-565 with file initial conditions (no h provided) */
int orc_sim_setup(orc_sim* s) {
  const int N = s->N;
  const orc_params* p = &s->p;
  const double hg = orc_h_guess(p, N, s->r);
  for (int i=0; i<N; i++) s->h[i] = hg;
  mac_snapshot(s);
  if (compute_h_masked(p, N, s->r, s->m, 1, s->h, s->rho, s->invomega, s->zeta, s->hfactor,
                       s->u, s->sound, s->pressure, NULL, sim_stars(s))) return -1;
  if (compute_h_masked(p, N, s->r, s->m, 0, s->h, s->rho, s->invomega, s->zeta, s->hfactor,
                       s->u, s->sound, s->pressure, NULL, sim_stars(s))) return -1;
  s->t = 0.0;
  s->n = 0; s->nresync = 0;
  for (int i=0; i<N; i++) { s->level[i] = 0; s->levelneib[i] = 0; s->ptlast[i] = s->t; }
  /* a relative opening criterion needs forces to start from: the first pass uses the geometric
     one, then the tree is stocked again with those forces and the forces recomputed, without a
     new h-rho pass (SphSimulation.cpp:380-470) */
  const int mac = s->p.gravity_mac;
  const int twice = p->self_gravity && !p->grav_direct && mac != 0;
  if (twice) s->p.gravity_mac = 0;
  if (sim_forces(s, NULL, 1)) return -1;
  if (twice) {
    s->p.gravity_mac = mac;
    mac_snapshot(s);
    if (sim_forces(s, NULL, 0)) return -1;
  }
  star_forces(s);             /* the initial N-body forces (SphSimulation.cpp:490-502) */
  star_end(s);
  if (p->nlevels > 1) {
    block_timesteps(s);
    end_timestep(s);          /* pdt = 0: sets r0, v0, a0, u0, dudt0 and the first dt */
  } else {
    s->timestep = orc_timesteps(p, N, s->h, s->sound, s->div_v, s->a, s->u, s->dudt, NULL);
    if (s->ns > 0) s->timestep = fmin(s->timestep, star_timestep(s));
    memcpy(s->r0, s->r, sizeof(double)*3*N);
    memcpy(s->v0, s->v, sizeof(double)*3*N);
    memcpy(s->a0, s->a, sizeof(double)*3*N);
    memcpy(s->u0, s->u, sizeof(double)*N);
    memcpy(s->dudt0, s->dudt, sizeof(double)*N);
  }
  s->tlast = s->t;
  return 0;
}

static void wrap_with_r0(orc_sim* s) {
  const orc_params* p = &s->p;
  if (!p->periodic) return;
  for (int i=0; i<s->N; i++)
    for (int k=0; k<3; k++) {
      double size = p->boxmax[k] - p->boxmin[k];
      if (s->r[3*i+k] < p->boxmin[k]) { s->r[3*i+k] += size; s->r0[3*i+k] += size; }
      if (s->r[3*i+k] > p->boxmax[k]) { s->r[3*i+k] -= size; s->r0[3*i+k] -= size; }
    }
}

/* SphSimulation::MainLoop (SphSimulation.cpp:574-880) with block timesteps */
static int sim_step_block(orc_sim* s) {
  const int N = s->N;
  const orc_params* p = &s->p;
  s->n = s->n + 1;
  s->t = s->t + s->timestep;
  /* SphLeapfrogKDK::AdvanceParticles (SphLeapfrogKDK.cpp:76-127) */
  for (int i=0; i<N; i++) {
    const double dt = s->t - s->ptlast[i];
    for (int k=0; k<3; k++)
      s->r[3*i+k] = s->r0[3*i+k] + s->v0[3*i+k]*dt + 0.5*s->a0[3*i+k]*dt*dt;
    for (int k=0; k<3; k++) s->v[3*i+k] = s->v0[3*i+k] + s->a0[3*i+k]*dt;
    if (p->energy_integration) s->u[i] = s->u0[i] + s->dudt0[i]*dt;
    s->active[i] = (s->n - s->nlast[i] == s->nstep[i]);
  }
  wrap_with_r0(s);
  int activecount;
  do {
    if (sim_forces(s, s->active, 1)) return -1;
    memset(s->active, 0, (size_t) N);
    activecount = check_timesteps(s);
  } while (activecount > 0);
  block_timesteps(s);
  end_timestep(s);
  s->tlast = s->t;
  return 0;
}

/* SphSimulation.cpp:574-880 */
int orc_sim_step(orc_sim* s) {
  const int N = s->N;
  const orc_params* p = &s->p;
  mac_snapshot(s);
  if (p->nlevels > 1) return sim_step_block(s);
  const double dtstep = s->timestep;
  s->t = s->t + s->timestep;
  const double dt = s->t - s->tlast;
  orc_advance(N, dt, p->energy_integration, s->r0, s->v0, s->a0, s->u0, s->dudt0, s->r, s->v,
              s->u);
  /* NbodyLeapfrogKDK::AdvanceParticles (NbodyLeapfrogKDK.cpp:253-289) */
  for (int i=0; i<3*s->ns; i++) {
    s->sr[i] = s->sr0[i] + s->sv0[i]*dt + 0.5*s->sa0[i]*dt*dt;
    s->sv[i] = s->sv0[i] + s->sa0[i]*dt;
  }
  wrap_with_r0(s);
  if (sim_forces(s, NULL, 1)) return -1;
  star_forces(s);
  /* NbodyLeapfrogKDK::CorrectionTerms (:296-326) */
  for (int i=0; i<3*s->ns; i++) s->sv[i] += 0.5*(s->sa[i] - s->sa0[i])*dt;
  double newstep = orc_timesteps(p, N, s->h, s->sound, s->div_v, s->a, s->u, s->dudt, NULL);
  if (s->ns > 0) newstep = fmin(newstep, star_timestep(s));
  orc_end_step(N, dtstep, p->energy_integration, s->a, s->dudt, s->v, s->u, s->r0, s->v0,
               s->a0, s->u0, s->dudt0, s->r);
  star_end(s);
  s->tlast = s->t;
  s->timestep = newstep;
  s->nfinished = N;
  return 0;
}

int orc_sim_get(orc_sim* s, const char* f, double* out) {
  const int N = s->N;
  struct { const char* name; double* p; int w; } tab[] = {
    {"r", s->r, 3}, {"v", s->v, 3}, {"a", s->a, 3}, {"atree", s->atree, 3}, {"m", s->m, 1},
    {"h", s->h, 1}, {"rho", s->rho, 1}, {"u", s->u, 1}, {"dudt", s->dudt, 1},
    {"div_v", s->div_v, 1}, {"gpot", s->gpot, 1}, {"invomega", s->invomega, 1},
    {"zeta", s->zeta, 1}, {"sound", s->sound, 1}, {"pressure", s->pressure, 1},
    {"hfactor", s->hfactor, 1}, {"tlast", s->ptlast, 1}, {"dt", s->pdt, 1}};
  for (unsigned k=0; k<sizeof(tab)/sizeof(tab[0]); k++)
    if (!strcmp(tab[k].name, f)) {
      memcpy(out, tab[k].p, sizeof(double)*N*tab[k].w);
      return 0;
    }
  struct { const char* name; double* p; int w; } stab[] = {
    {"star_r", s->sr, 3}, {"star_v", s->sv, 3}, {"star_a", s->sa, 3}, {"star_adot", s->sadot, 3},
    {"star_gpot", s->sgpot, 1}};
  for (unsigned k=0; k<sizeof(stab)/sizeof(stab[0]); k++)
    if (!strcmp(stab[k].name, f)) {
      if (s->ns > 0) memcpy(out, stab[k].p, sizeof(double)*s->ns*stab[k].w);
      return 0;
    }
  struct { const char* name; int* p; } itab[] = {
    {"level", s->level}, {"levelneib", s->levelneib}, {"nstep", s->nstep}, {"nlast", s->nlast}};
  for (unsigned k=0; k<sizeof(itab)/sizeof(itab[0]); k++)
    if (!strcmp(itab[k].name, f)) {
      for (int i=0; i<N; i++) out[i] = (double) itab[k].p[i];
      return 0;
    }
  return -1;
}
double orc_sim_time(orc_sim* s) { return s->t; }
double orc_sim_timestep(orc_sim* s) { return s->timestep; }

/* n, level_max, level_step, nresync, dt_max, particles that ended a step in the last step */
void orc_sim_block_state(orc_sim* s, int* iout, double* dout) {
  iout[0] = s->n; iout[1] = s->level_max; iout[2] = s->level_step; iout[3] = s->nresync;
  dout[0] = s->dt_max; dout[1] = (double) s->nfinished;
}
