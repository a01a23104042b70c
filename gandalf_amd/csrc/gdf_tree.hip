// Device tree: Morton keys, radix sort, LBVH topology, bottom-up cell moments, walk groups.
//
// Replaces KDTree::BuildTree / StockTree (src/Tree/KDTree.cpp:220-313, :760-1083) behind
// HydroTree::BuildTree (src/Tree/HydroTree.cpp:310-372).  Instead of median QuickSelect splits
// the particles are sorted along a 63-bit Morton curve and the binary radix tree over the sorted
// keys (Karras 2012) supplies the cells.  Cell moments follow the reference's bottom-up stock:
// mass, centre of mass, traceless quadrupole with the parallel-axis merge (KDTree.cpp:950-1051),
// bounding box, hmax and the geometric-MAC distance cdistsqd = max(rmax^2, hmax^2)/theta^2
// (KDTree.cpp:917,997).  The particle arrays are physically permuted into key order so every
// later pass streams coalesced SoA data.
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>

#include "gdf_internal.h"
#include "gdf_device.h"

namespace gdf {
namespace {

constexpr int BS = 256;

__global__ void k_bounds_partial(int N, const double* __restrict__ x, const double* __restrict__ y,
                                 const double* __restrict__ z, double* __restrict__ part) {
  __shared__ double s[6][BS];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = blockIdx.x*BS + threadIdx.x; i < N; i += gridDim.x*BS) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], v[k]); hi[k] = fmax(hi[k], v[k]); }
  }
  for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = lo[k]; s[3+k][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        s[3+k][threadIdx.x] = fmax(s[3+k][threadIdx.x], s[3+k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x*6 + threadIdx.x] = s[threadIdx.x][0];
}

// out: lo[3], scale; one block of BS threads folds the per-block partials
__global__ __launch_bounds__(BS) void k_bounds_final(int nb, const double* __restrict__ part,
                                                     double* __restrict__ out) {
  __shared__ double s[6][BS];
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int b = threadIdx.x; b < nb; b += BS)
    for (int k = 0; k < 3; k++) {
      lo[k] = fmin(lo[k], part[b*6 + k]);
      hi[k] = fmax(hi[k], part[b*6 + 3 + k]);
    }
  for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = lo[k]; s[3+k][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; k++) {
        s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + w]);
        s[3+k][threadIdx.x] = fmax(s[3+k][threadIdx.x], s[3+k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  double span = s[3][0] - s[0][0];
  if (s[4][0] - s[1][0] > span) span = s[4][0] - s[1][0];
  if (s[5][0] - s[2][0] > span) span = s[5][0] - s[2][0];
  if (!(span > 0.0)) span = 1.0;
  out[0] = s[0][0]; out[1] = s[1][0]; out[2] = s[2][0];
  out[3] = 2097152.0/span;
}

// ---- serving batches: per-segment bounds and segment-prefixed keys ----
// partial-reduction blocks per segment: kMaxSeg x 16 x 6 partials fit the reduction buffer
constexpr int kSegBlocks = 16;
constexpr int kSelBlock = 256;      // workgroup of the flag selection; 16 flags per thread

// grid (kSegBlocks, nseg): block (b, s) folds a strided share of segment s
__global__ void k_seg_bounds_partial(const int* __restrict__ off, const double* __restrict__ x,
                                     const double* __restrict__ y, const double* __restrict__ z,
                                     double* __restrict__ part) {
  __shared__ double sh[6][BS];
  const int sg = blockIdx.y, b = blockIdx.x;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = off[sg] + b*BS + threadIdx.x; i < off[sg + 1]; i += kSegBlocks*BS) {
    const double v[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; k++) { lo[k] = fmin(lo[k], v[k]); hi[k] = fmax(hi[k], v[k]); }
  }
  for (int k = 0; k < 3; k++) { sh[k][threadIdx.x] = lo[k]; sh[3+k][threadIdx.x] = hi[k]; }
  __syncthreads();
  for (int w = BS/2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; k++) {
        sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + w]);
        sh[3+k][threadIdx.x] = fmax(sh[3+k][threadIdx.x], sh[3+k][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[(sg*kSegBlocks + b)*6 + threadIdx.x] = sh[threadIdx.x][0];
}

// one 64-thread block per segment: lower corner and scale of a 2^19 grid per axis
__global__ void k_seg_bounds_final(const double* __restrict__ part, double* __restrict__ bnd) {
  const int sg = blockIdx.x, t = threadIdx.x;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  if (t < kSegBlocks)
    for (int k = 0; k < 3; k++) {
      lo[k] = part[(sg*kSegBlocks + t)*6 + k];
      hi[k] = part[(sg*kSegBlocks + t)*6 + 3 + k];
    }
  for (int k = 0; k < 3; k++)
    for (int d = 32; d >= 1; d >>= 1) {
      lo[k] = fmin(lo[k], __shfl_xor(lo[k], d, 64));
      hi[k] = fmax(hi[k], __shfl_xor(hi[k], d, 64));
    }
  if (t != 0) return;
  double span = fmax(fmax(hi[0] - lo[0], hi[1] - lo[1]), hi[2] - lo[2]);
  if (!(span > 0.0)) span = 1.0;
  bnd[4*sg] = lo[0]; bnd[4*sg + 1] = lo[1]; bnd[4*sg + 2] = lo[2];
  bnd[4*sg + 3] = 524288.0/span;
}

__device__ __forceinline__ uint64_t spread3(uint64_t x) {
  x &= 0x1fffffULL;
  x = (x | (x << 32)) & 0x1f00000000ffffULL;
  x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
  x = (x | (x << 8))  & 0x100f00f00f00f00fULL;
  x = (x | (x << 4))  & 0x10c30c30c30c30c3ULL;
  x = (x | (x << 2))  & 0x1249249249249249ULL;
  return x;
}

__global__ void k_morton(int N, const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, const double* __restrict__ bnd,
                         uint64_t* __restrict__ keys, int* __restrict__ vals) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  const double scale = bnd[3];
  const double v[3] = {x[i], y[i], z[i]};
  uint64_t q[3];
  for (int k = 0; k < 3; k++) {
    const double f = __dmul_rn(__dadd_rn(v[k], -bnd[k]), scale);
    q[k] = f >= 2097151.0 ? 2097151ULL : (uint64_t) f;
  }
  keys[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  vals[i] = i;
}

struct PermArgs {
  const double* src[F_NPERSIST]; double* dst[F_NPERSIST];
  const int* isrc[I_NINT]; int* idst[I_NINT];
};

__global__ void k_permute(int N, const int* __restrict__ vals, PermArgs a,
                          const int* __restrict__ io, int* __restrict__ io_alt) {
  const int s = blockIdx.x*BS + threadIdx.x;
  if (s >= N) return;
  const int o = vals[s];
  #pragma unroll
  for (int f = 0; f < F_NPERSIST; f++) a.dst[f][s] = a.src[f][o];
  #pragma unroll
  for (int f = 0; f < I_NINT; f++) a.idst[f][s] = a.isrc[f][o];
  io_alt[s] = io[o];
}

// block timesteps: the derived fields an inactive particle keeps until its step ends (its
// neighbours read rho, P, Omega, zeta, c_s, hfactor; diagnostics read gpot; a, atree, du/dt and
// div v are its visible state); the search radius, the timestep criterion, the N-body jerk and
// the gravity walk's partial sums are rewritten before every use and are not moved
constexpr int kPermD[] = {D_AX, D_AY, D_AZ, D_ATX, D_ATY, D_ATZ, D_RHO, D_DUDT, D_DIVV, D_GPOT,
                          D_INVOMEGA, D_ZETA, D_SOUND, D_PRESSURE, D_HFACTOR};
constexpr int kNPermD = sizeof(kPermD)/sizeof(kPermD[0]);
struct PermDArgs { const double* src[kNPermD]; double* dst[kNPermD]; };

__global__ void k_permute_d(int N, const int* __restrict__ vals, PermDArgs a) {
  const int s = blockIdx.x*BS + threadIdx.x;
  if (s >= N) return;
  const int o = vals[s];
  #pragma unroll
  for (int f = 0; f < kNPermD; f++) a.dst[f][s] = a.src[f][o];
}

__device__ __forceinline__ int delta(const uint64_t* __restrict__ key, int N, int i, long long j) {
  if (j < 0 || j > N - 1) return -1;
  const uint64_t a = key[i], b = key[j];
  if (a != b) return __clzll((long long) (a ^ b));
  return 64 + __clz((int) (i ^ (int) j));
}

// segment-prefixed key: segment id in bits 57-63, 19 bits per axis below
__global__ void k_morton_seg(int N, const double* __restrict__ x, const double* __restrict__ y,
                             const double* __restrict__ z, const int* __restrict__ seg,
                             const double* __restrict__ bnd, uint64_t* __restrict__ keys,
                             int* __restrict__ vals) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N) return;
  const int sg = seg[i];
  const double scale = bnd[4*sg + 3];
  const double v[3] = {x[i], y[i], z[i]};
  uint64_t q[3];
  for (int k = 0; k < 3; k++) {
    const double f = __dmul_rn(__dadd_rn(v[k], -bnd[4*sg + k]), scale);
    q[k] = f >= 524287.0 ? 524287ULL : (uint64_t) f;
  }
  keys[i] = ((uint64_t) sg << 57) | (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
  vals[i] = i;
}

// the node spanning exactly segment s is its subtree root (the keys of a segment share the
// segment prefix and no other key does)
__global__ void k_seg_roots(int N, int nseg, const int* __restrict__ off,
                            const int* __restrict__ first, const int* __restrict__ last,
                            const int* __restrict__ seg, int* __restrict__ root) {
  const int n = blockIdx.x*BS + threadIdx.x;
  if (n < N - 1) {
    const int f = first[n], l = last[n], sg = seg[f];
    if (f == off[sg] && l == off[sg + 1] - 1) root[sg] = n;
  } else if (n < 2*N - 1) {
    const int leaf = n - (N - 1), sg = seg[leaf];
    if (off[sg + 1] - off[sg] == 1) root[sg] = n;
  }
}

// The walking group's factor of the gadget2 / eigenmac opening criterion (the reference's
// walking leaf cell's amin / macfactor, stocked from its particles, KDTree.cpp:900-903):
// gadget2 the smallest |atree|, eigenmac the largest gpot^(-2/3).  One wavefront per group;
// perm (non-NULL) maps a group's sorted slots to the force arrays' order when the build did not
// permute them.
__global__ __launch_bounds__(BS) void k_group_mac(int ngroups, const int* __restrict__ grp_first,
                                                  const int* __restrict__ grp_count,
                                                  const int* __restrict__ perm,
                                                  const double* __restrict__ ax,
                                                  const double* __restrict__ ay,
                                                  const double* __restrict__ az,
                                                  const double* __restrict__ gpot, int mode,
                                                  double* __restrict__ out) {
#pragma clang fp contract(off)
  const int g = blockIdx.x*(BS/kWave) + (int) threadIdx.x/kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (g >= ngroups) return;
  const bool gadget = mode == GDF_MAC_GADGET2;
  double v = gadget ? 9.9e30 : 0.0;
  if (lane < grp_count[g]) {
    int i = grp_first[g] + lane;
    if (perm) i = perm[i];
    v = gadget ? sqrt(ax[i]*ax[i] + ay[i]*ay[i] + az[i]*az[i])
               : pow(gpot[i], -0.66666666666666666666666);
  }
  for (int o = 32; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, kWave);
    v = gadget ? fmin(v, w) : fmax(v, w);
  }
  if (lane == 0) out[g] = v;
}

int group_mac(gdf_ctx* c, const int* perm) {
  if (c->phys.gravity_mac == GDF_MAC_GEOMETRIC || c->ngroups <= 0) return GDF_OK;
  constexpr int gpb = BS/kWave;
  k_group_mac<<<(c->ngroups + gpb - 1)/gpb, BS, 0, c->stream>>>(
      c->ngroups, c->grp_first, c->grp_count, perm, c->df[D_ATX], c->df[D_ATY], c->df[D_ATZ],
      c->df[D_GPOT], c->phys.gravity_mac, c->grp_gmac);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

__global__ void k_grp_roots(int ngroups, const int* __restrict__ grp_first,
                            const int* __restrict__ seg, const int* __restrict__ root,
                            int* __restrict__ grp_root) {
  const int g = blockIdx.x*BS + threadIdx.x;
  if (g < ngroups) grp_root[g] = root[seg[grp_first[g]]];
}

// Karras 2012, "Maximizing parallelism in the construction of BVHs, octrees, and k-d trees"
__global__ void k_lbvh(int N, const uint64_t* __restrict__ key, int* __restrict__ left,
                       int* __restrict__ right, int* __restrict__ first, int* __restrict__ last,
                       int* __restrict__ parent, int4* __restrict__ lnk) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i >= N - 1) return;
  const int d = (delta(key, N, i, (long long) i + 1) - delta(key, N, i, (long long) i - 1)) >= 0 ? 1 : -1;
  const int dmin = delta(key, N, i, (long long) i - d);
  long long lmax = 2;
  while (delta(key, N, i, (long long) i + lmax*d) > dmin) lmax *= 2;
  long long l = 0;
  for (long long t = lmax/2; t >= 1; t /= 2)
    if (delta(key, N, i, (long long) i + (l + t)*d) > dmin) l += t;
  const long long j = (long long) i + l*d;
  const int dnode = delta(key, N, i, j);
  long long s = 0, div = 2, step;
  do {
    step = (l + div - 1)/div;
    if (delta(key, N, i, (long long) i + (s + step)*d) > dnode) s += step;
    div *= 2;
  } while (step > 1);
  const int gamma = (int) ((long long) i + s*d + (d < 0 ? d : 0));
  const int f = (int) (i < j ? i : j), la = (int) (i < j ? j : i);
  first[i] = f; last[i] = la;
  const int L = (f == gamma) ? (N - 1 + gamma) : gamma;
  const int R = (la == gamma + 1) ? (N - 1 + gamma + 1) : gamma + 1;
  left[i] = L; right[i] = R;
  lnk[i] = make_int4(f, la, L, R);
  parent[L] = i; parent[R] = i;
  if (i == 0) parent[0] = -1;
}

struct NodeArgs {
  double* bbmin; double* bbmax; double4* geo_c; double4* geo_h; double* mp;
};

__global__ void k_leaves(int N, const double* __restrict__ x, const double* __restrict__ y,
                         const double* __restrict__ z, const double* __restrict__ m,
                         const double* __restrict__ h, double theta2, NodeArgs a,
                         int* __restrict__ visit, int* __restrict__ parent) {
  const int l = blockIdx.x*BS + threadIdx.x;
  if (l >= N) return;
  if (l < N - 1) visit[l] = 0;
  if (N == 1) parent[0] = -1;
  const int n = N - 1 + l;
  const double r[3] = {x[l], y[l], z[l]};
  for (int k = 0; k < 3; k++) { a.bbmin[3*n+k] = r[k]; a.bbmax[3*n+k] = r[k]; a.mp[10*n+k] = r[k]; }
  a.mp[10*n+3] = m[l];
  for (int k = 4; k < 10; k++) a.mp[10*n+k] = 0.0;
  const double hh = h[l];
  a.geo_c[n] = make_double4(r[0], r[1], r[2], hh);
  a.geo_h[n] = make_double4(0.0, 0.0, 0.0, __ddiv_rn(__dmul_rn(hh, hh), theta2));
}

// Moments of one node as a refit climb carries them: bounding box, centre of mass, mass,
// quadrupole moment and the largest smoothing length below the node.
struct NodeD {
  double lo[3], hi[3], com[3], m, q[5], hmax;
};

__device__ __forceinline__ void load_node(int n, const NodeArgs& a, NodeD& d) {
  for (int k = 0; k < 3; k++) {
    d.lo[k] = a.bbmin[3*n+k]; d.hi[k] = a.bbmax[3*n+k]; d.com[k] = a.mp[10*n+k];
  }
  d.m = a.mp[10*n+3];
  for (int k = 0; k < 5; k++) d.q[k] = a.mp[10*n+4+k];
  d.hmax = a.geo_c[n].w;
}

// merge children L (left) and R (right) into o; operation order of KDTree.cpp:950-1051
__device__ __forceinline__ void merge_d(const NodeD& L, const NodeD& R, NodeD& o) {
#pragma clang fp contract(off)
  const double m = L.m + R.m;
  for (int k = 0; k < 3; k++) {
    o.com[k] = m > 0 ? (L.m*L.com[k] + R.m*R.com[k])/m : 0.0;
    o.lo[k] = fmin(L.lo[k], R.lo[k]);
    o.hi[k] = fmax(L.hi[k], R.hi[k]);
  }
  o.m = m;
  o.hmax = fmax(L.hmax, R.hmax);
  for (int k = 0; k < 5; k++) o.q[k] = 0.0;
  for (int c = 0; c < 2; c++) {
    const NodeD& C = c == 0 ? L : R;
    const double mi = C.m;
    if (!(mi > 0)) continue;
    double dr[3];
    for (int k = 0; k < 3; k++) dr[k] = C.com[k] - o.com[k];
    const double drsqd = dr[0]*dr[0] + dr[1]*dr[1] + dr[2]*dr[2];
    for (int k = 0; k < 5; k++) o.q[k] += C.q[k];
    o.q[0] += mi*(3.0*dr[0]*dr[0] - drsqd);
    o.q[1] += mi*3.0*dr[0]*dr[1];
    o.q[2] += mi*(3.0*dr[1]*dr[1] - drsqd);
    o.q[3] += mi*3.0*dr[2]*dr[0];
    o.q[4] += mi*3.0*dr[2]*dr[1];
  }
}

__device__ __forceinline__ void store_node(int n, const NodeD& d, double theta2,
                                           const NodeArgs& a) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; k++) {
    a.mp[10*n+k] = d.com[k]; a.bbmin[3*n+k] = d.lo[k]; a.bbmax[3*n+k] = d.hi[k];
  }
  a.mp[10*n+3] = d.m;
  for (int k = 0; k < 5; k++) a.mp[10*n+4+k] = d.q[k];
  double c[3], hf[3];
  for (int k = 0; k < 3; k++) { c[k] = 0.5*(d.lo[k] + d.hi[k]); hf[k] = 0.5*(d.hi[k] - d.lo[k]); }
  const double rmaxsqd = hf[0]*hf[0] + hf[1]*hf[1] + hf[2]*hf[2];
  a.geo_c[n] = make_double4(c[0], c[1], c[2], d.hmax);
  a.geo_h[n] = make_double4(hf[0], hf[1], hf[2], fmax(rmaxsqd, d.hmax*d.hmax)/theta2);
}

// new smoothing-length bound of internal node n whose box is unchanged
__device__ __forceinline__ void store_h(int n, double hmax, double theta2, const NodeArgs& a) {
#pragma clang fp contract(off)
  double4 gc = a.geo_c[n];
  double4 gh = a.geo_h[n];
  const double rmaxsqd = gh.x*gh.x + gh.y*gh.y + gh.z*gh.z;
  gc.w = hmax;
  gh.w = fmax(rmaxsqd, hmax*hmax)/theta2;
  a.geo_c[n] = gc;
  a.geo_h[n] = gh;
}


// Bottom-up refit in three phases.  The second child to reach a node merges it (an arrival
// counter per node decides which), so every node is merged exactly once from finished children
// in the fixed (left, right) operand order.  Synchronising through device-scope atomics would
// write back and invalidate the per-XCD L2 on every tree level, so the counters stay within one
// workgroup: phase 1 gives each workgroup a chunk of RB consecutive leaves and merges the nodes
// whose particle range lies inside the chunk (an internal node is one end of its own range, so
// its counter fits the chunk's LDS); the maximal such nodes form the chunk's frontier.  Phase 2
// repeats this over RS chunks per workgroup starting from those frontiers, and phase 3 is one
// workgroup that climbs the few remaining nodes to the root.
constexpr int RB = 1024;   // leaves per phase-1 workgroup
constexpr int RS = 32;     // phase-1 chunks per phase-2 workgroup

// Phase 1, fused with the leaf moments.  The chunk's topology is staged in LDS, and each
// climbing thread carries its node's moments in registers: before it counts its arrival at a
// parent it leaves them in its own LDS slot, and the exchange that counts the arrival hands the
// second arriver the first one's slot, so a merge reads nothing from global memory.  A slot is
// read only after its owner has stopped climbing, hence is never overwritten while still needed.
template <bool HONLY>
__global__ __launch_bounds__(RB) void k_refit_chunks(
    int N, const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, const double* __restrict__ mass, const double* __restrict__ h,
    const int* __restrict__ parent, const int* __restrict__ left, const int* __restrict__ first,
    const int* __restrict__ last, double theta2, NodeArgs a, int* __restrict__ visit,
    int* __restrict__ front1, int* __restrict__ cnt1, int* __restrict__ front2_count) {
#pragma clang fp contract(off)
  constexpr int ND = HONLY ? 1 : 16;    // doubles carried per node
  __shared__ double sd[ND*RB];
  __shared__ int arr[RB];               // 0, or 1 + slot of the first arrival
  __shared__ int lpar[RB];              // parent of leaf c0 + t
  __shared__ int ipar[RB];              // parent of internal node c0 + t
  __shared__ int lft[RB];               // left child of internal node c0 + t, -1 if not inside
  __shared__ int nfront;
  const int t = threadIdx.x;
  const int c0 = blockIdx.x*RB;
  const int c1 = min(c0 + RB, N) - 1;
  const int l = c0 + t;
  arr[t] = 0;
  if (t == 0) nfront = 0;
  if (blockIdx.x == 0 && t == 0) *front2_count = 0;   // phase 2 counts its frontier here
  NodeD d;
  if (l <= c1) {
    lpar[t] = parent[N - 1 + l];
    if (l < N - 1) {
      ipar[t] = parent[l];
      lft[t] = (first[l] >= c0 && last[l] <= c1) ? left[l] : -1;
      visit[l] = 0;
    } else {
      lft[t] = -1;
    }
    const int n = N - 1 + l;
    const double hh = h[l];
    d.hmax = hh;
    if (HONLY) {
      double4 gc = a.geo_c[n]; gc.w = hh; a.geo_c[n] = gc;
      double4 gh = a.geo_h[n]; gh.w = hh*hh/theta2; a.geo_h[n] = gh;
    } else {
      const double r[3] = {x[l], y[l], z[l]};
      for (int k = 0; k < 3; k++) { d.lo[k] = r[k]; d.hi[k] = r[k]; d.com[k] = r[k]; }
      d.m = mass[l];
      for (int k = 0; k < 5; k++) d.q[k] = 0.0;
      for (int k = 0; k < 3; k++) { a.bbmin[3*n+k] = r[k]; a.bbmax[3*n+k] = r[k]; a.mp[10*n+k] = r[k]; }
      a.mp[10*n+3] = d.m;
      for (int k = 4; k < 10; k++) a.mp[10*n+k] = 0.0;
      a.geo_c[n] = make_double4(r[0], r[1], r[2], hh);
      a.geo_h[n] = make_double4(0.0, 0.0, 0.0, hh*hh/theta2);
    }
  }
  __syncthreads();
  if (l <= c1) {
    int n = N - 1 + l;
    int p = lpar[t];
    while (p >= 0) {
      const int lf = (p >= c0 && p <= c1) ? lft[p - c0] : -1;
      if (lf < 0) {
        // n is a maximal node of this chunk: at most RB of them
        front1[c0 + atomicAdd(&nfront, 1)] = n;
        break;
      }
      if (HONLY) {
        sd[t] = d.hmax;
      } else {
        for (int k = 0; k < 3; k++) {
          sd[k*RB + t] = d.lo[k]; sd[(3 + k)*RB + t] = d.hi[k]; sd[(6 + k)*RB + t] = d.com[k];
        }
        sd[9*RB + t] = d.m;
        for (int k = 0; k < 5; k++) sd[(10 + k)*RB + t] = d.q[k];
        sd[15*RB + t] = d.hmax;
      }
      const int old = __hip_atomic_exchange(&arr[p - c0], t + 1, __ATOMIC_ACQ_REL,
                                            __HIP_MEMORY_SCOPE_WORKGROUP);
      if (old == 0) break;
      const int o = old - 1;
      if (HONLY) {
        d.hmax = fmax(d.hmax, sd[o]);
        store_h(p, d.hmax, theta2, a);
      } else {
        NodeD e, r;
        for (int k = 0; k < 3; k++) {
          e.lo[k] = sd[k*RB + o]; e.hi[k] = sd[(3 + k)*RB + o]; e.com[k] = sd[(6 + k)*RB + o];
        }
        e.m = sd[9*RB + o];
        for (int k = 0; k < 5; k++) e.q[k] = sd[(10 + k)*RB + o];
        e.hmax = sd[15*RB + o];
        if (lf == n) merge_d(d, e, r);
        else merge_d(e, d, r);
        store_node(p, r, theta2, a);
        d = r;
      }
      n = p;
      p = ipar[p - c0];
    }
  }
  __syncthreads();
  if (t == 0) cnt1[blockIdx.x] = nfront;
}

// Topology of the next node of a climb, loaded together with the moments of the current
// merge so that each level of phases 2 and 3 waits on global memory about once.
struct Topo {
  int first, last, left, right, parent;
  double rmax2;   // squared half-diagonal of the node's box (smoothing-length refit only)
};

template <bool HONLY>
__device__ __forceinline__ void load_topo(int p, const int* __restrict__ parent,
                                          const int* __restrict__ left,
                                          const int* __restrict__ right,
                                          const int* __restrict__ first,
                                          const int* __restrict__ last, const NodeArgs& a,
                                          Topo& q) {
#pragma clang fp contract(off)
  q.first = first[p]; q.last = last[p]; q.left = left[p]; q.right = right[p];
  q.parent = parent[p];
  if (HONLY) {
    // the box only: another thread may be writing the node's .w bound
    const double hx = a.geo_h[p].x, hy = a.geo_h[p].y, hz = a.geo_h[p].z;
    q.rmax2 = hx*hx + hy*hy + hz*hz;
  }
}

// Climb from node n through the nodes whose leaf range lies in [lo, hi] (phase 2) or up to
// the root (phase 3, lo = 0 and hi = N - 1).  The second arrival at a node, counted in
// cnt[p - cbase], merges it from its own moments (registers) and its sibling's (global memory,
// written before the sibling's thread counted its arrival).  A node whose parent leaves the
// range goes to the next frontier.
template <bool HONLY>
__device__ __forceinline__ void climb(int n, int lo, int hi, int* cnt, int cbase,
                                      const int* __restrict__ parent,
                                      const int* __restrict__ left,
                                      const int* __restrict__ right,
                                      const int* __restrict__ first,
                                      const int* __restrict__ last, double theta2,
                                      const NodeArgs& a, int* __restrict__ front_next,
                                      int* __restrict__ front_count) {
#pragma clang fp contract(off)
  NodeD d;
  if (HONLY) d.hmax = a.geo_c[n].w;
  else load_node(n, a, d);
  int p = parent[n];
  Topo q;
  if (p >= 0) load_topo<HONLY>(p, parent, left, right, first, last, a, q);
  while (p >= 0) {
    if (q.first < lo || q.last > hi) {
      front_next[atomicAdd(front_count, 1)] = n;
      break;
    }
    const int old = __hip_atomic_fetch_add(&cnt[p - cbase], 1, __ATOMIC_ACQ_REL,
                                           __HIP_MEMORY_SCOPE_WORKGROUP);
    if (old == 0) break;
    const bool isleft = q.left == n;
    const int sib = isleft ? q.right : q.left;
    const int pp = q.parent;
    Topo qn;
    if (pp >= 0) load_topo<HONLY>(pp, parent, left, right, first, last, a, qn);
    if (HONLY) {
      d.hmax = fmax(d.hmax, a.geo_c[sib].w);
      a.geo_c[p].w = d.hmax;
      a.geo_h[p].w = fmax(q.rmax2, d.hmax*d.hmax)/theta2;
    } else {
      NodeD e, r;
      load_node(sib, a, e);
      if (isleft) merge_d(d, e, r);
      else merge_d(e, d, r);
      store_node(p, r, theta2, a);
      d = r;
    }
    n = p;
    p = pp;
    q = qn;
  }
}

// Phase 2: one workgroup per RS consecutive phase-1 chunks climbs from their frontier nodes
// through the nodes inside its leaf range.  Every arrival at such a node comes from this
// workgroup, and the node's index lies in the range, so the arrival counters live in LDS.
template <bool HONLY>
__global__ __launch_bounds__(RB) void k_refit_mid(
    int N, int nchunks, const int* __restrict__ parent, const int* __restrict__ left,
    const int* __restrict__ right, const int* __restrict__ first, const int* __restrict__ last,
    double theta2, NodeArgs a, const int* __restrict__ front1, const int* __restrict__ cnt1,
    int* __restrict__ front2) {
  __shared__ int arr[RB*RS];
  __shared__ int pre[RS + 1];
  const int k0 = blockIdx.x*RS, k1 = min(k0 + RS, nchunks);
  const int lo = k0*RB, hi = min(k1*RB, N) - 1;
  for (int i = threadIdx.x; i < RB*RS; i += RB) arr[i] = 0;
  if (threadIdx.x == 0) {
    pre[0] = 0;
    for (int c = 0; c < RS; c++) pre[c + 1] = pre[c] + (k0 + c < k1 ? cnt1[k0 + c] : 0);
  }
  __syncthreads();
  const int total = pre[RS];
  for (int e = threadIdx.x; e < total; e += RB) {
    int c = 0;
    while (pre[c + 1] <= e) c++;
    climb<HONLY>(front1[(k0 + c)*RB + (e - pre[c])], lo, hi, arr, lo, parent, left, right,
                 first, last, theta2, a, front2, &front2[N]);
  }
}

// Phase 3: one workgroup climbs from the phase-3 frontier to the root, counting arrivals in
// global memory (the nodes above the frontier may have any index).
template <bool HONLY>
__global__ __launch_bounds__(RB) void k_refit_top(
    int N, const int* __restrict__ parent, const int* __restrict__ left,
    const int* __restrict__ right, const int* __restrict__ first, const int* __restrict__ last,
    double theta2, NodeArgs a, const int* __restrict__ front2, int* __restrict__ visit) {
  const int nf = front2[N];
  // every range lies in [0, N - 1], so the climb never reaches its frontier branch
  for (int k = threadIdx.x; k < nf; k += RB)
    climb<HONLY>(front2[k], 0, N - 1, visit, 0, parent, left, right, first, last, theta2, a,
                 nullptr, nullptr);
}

template <bool HONLY>
void launch_refit(gdf_ctx* c, const NodeArgs& na, double theta2) {
  const int N = c->N;
  hipStream_t s = c->stream;
  const int nch = (N + RB - 1)/RB;
  int* front1 = c->front;                          // nch*RB entries
  int* cnt1 = front1 + (size_t) nch*RB;            // nch counts
  int* front2 = cnt1 + nch;                        // N entries + count
  k_refit_chunks<HONLY><<<nch, RB, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z], c->pf[F_M],
                                           c->pf[F_H], c->parent, c->left, c->first, c->last,
                                           theta2, na, c->visit, front1, cnt1, front2 + N);
  if (nch > 1) {
    const int nmid = (nch + RS - 1)/RS;
    k_refit_mid<HONLY><<<nmid, RB, 0, s>>>(N, nch, c->parent, c->left, c->right, c->first,
                                           c->last, theta2, na, front1, cnt1, front2);
    if (nmid > 1)
      k_refit_top<HONLY><<<1, RB, 0, s>>>(N, c->parent, c->left, c->right, c->first, c->last,
                                          theta2, na, front2, c->visit);
  }
}

__global__ void k_hleaves(int N, const double* __restrict__ h, double theta2, NodeArgs a,
                          int* __restrict__ visit) {
  const int l = blockIdx.x*BS + threadIdx.x;
  if (l >= N) return;
  if (l < N - 1) visit[l] = 0;
  const int n = N - 1 + l;
  const double hh = h[l];
  double4 gc = a.geo_c[n]; gc.w = hh; a.geo_c[n] = gc;
  double4 gh = a.geo_h[n]; gh.w = __ddiv_rn(__dmul_rn(hh, hh), theta2); a.geo_h[n] = gh;
}

// a node is a walk group iff it holds <= gmax particles and its parent holds more
__global__ void k_groups_mark(int N, int gmax, const int* __restrict__ parent,
                              const int* __restrict__ first, const int* __restrict__ last,
                              int* __restrict__ gstart, unsigned char* __restrict__ gflag) {
  const int n = blockIdx.x*BS + threadIdx.x;
  if (n >= 2*N - 1) return;
  const int cnt = n >= N - 1 ? 1 : last[n] - first[n] + 1;
  const int fst = n >= N - 1 ? n - (N - 1) : first[n];
  if (cnt > gmax) return;
  const int p = parent[n];
  const int pcnt = p < 0 ? 0x7fffffff : last[p] - first[p] + 1;
  if (p >= 0 && pcnt <= gmax) return;
  gstart[fst] = n;
  gflag[fst] = 1;
}

__global__ void k_groups_fill(int N, const int* __restrict__ nsel, const int* __restrict__ slots,
                              const int* __restrict__ gstart, const int* __restrict__ first,
                              const int* __restrict__ last, int* __restrict__ gnode,
                              int* __restrict__ gcount) {
  const int g = blockIdx.x*BS + threadIdx.x;
  if (g >= *nsel) return;
  const int s = slots[g];
  const int n = gstart[s];
  gnode[g] = n;
  gcount[g] = n >= N - 1 ? 1 : last[n] - first[n] + 1;
}

// Order-preserving compaction of n byte flags (0 or 1) into the indices of the set ones
// (count at out[cap]) by a single wavefront without LDS.  It may run while another stream's
// kernels fill the device: a look-back scan then stalls until its predecessor workgroups find
// a compute unit, and any workgroup that needs LDS waits until enough LDS is free on one CU
// (both measured at up to 2-4 ms beside the gravity walk); one LDS-free wavefront starts at
// once.  Each lane takes a contiguous run of 16-byte blocks (the flag buffer is zero-padded to
// a multiple of 16 bytes).
__global__ __launch_bounds__(64) void k_compact_flags(int n, const uint4* __restrict__ blocks,
                                                      int* __restrict__ out, int cap) {
  const int lane = threadIdx.x;
  const int nblk = (n + 15)/16;
  const int per = (nblk + kWave - 1)/kWave;
  const int b0 = min(lane*per, nblk), b1 = min(b0 + per, nblk);
  int cnt = 0;
  for (int b = b0; b < b1; b++) {
    const uint4 x = blocks[b];
    cnt += __popc(x.x) + __popc(x.y) + __popc(x.z) + __popc(x.w);   // bytes are 0 or 1
  }
  const int incl = wave_incl_scan(cnt);
  int o = incl - cnt;
  for (int b = b0; b < b1; b++) {
    const uint4 x = blocks[b];
    const unsigned wd[4] = {x.x, x.y, x.z, x.w};
    for (int k = 0; k < 4; k++) {
      if (!wd[k]) continue;
      for (int q = 0; q < 4; q++)
        if ((wd[k] >> (8*q)) & 1u) out[o++] = 16*b + 4*k + q;
    }
  }
  if (lane == kWave - 1) out[cap] = incl;
}

// one wavefront per walk group: does any of its particles have status 1?
__global__ void k_group_flags(int ngroups, const int* __restrict__ grp_first,
                              const int* __restrict__ grp_count, const int* __restrict__ status,
                              unsigned char* __restrict__ flags, int* __restrict__ ncand,
                              int* __restrict__ gnc, int* __restrict__ gnn,
                              int* __restrict__ zero, int nzero) {
  const int g = blockIdx.x*(BS/kWave) + threadIdx.x/kWave;
  const int lane = threadIdx.x % kWave;
  if (blockIdx.x == 0 && (int) threadIdx.x < nzero) zero[threadIdx.x] = 0;
  if (g >= ngroups) return;
  const bool on = lane < grp_count[g] && status[grp_first[g] + lane] == 1;
  const bool any = __ballot(on) != 0;
  if (lane == 0) {
    flags[g] = any ? 1 : 0;
    // the compaction reads whole 16-byte blocks: the bytes after the last flag read 0
    if (g == ngroups - 1)
      for (int k = ngroups; k & 15; k++) flags[k] = 0;
    if (!any) {
      if (ncand) ncand[g] = 0;
      if (gnc) { gnc[g] = 0; gnn[g] = 0; }
    }
  }
}

}  // namespace

TreeView tree_view(const gdf_ctx* c) {
  TreeView v;
  v.N = c->N; v.left = c->left; v.right = c->right; v.first = c->first; v.last = c->last;
  v.geo_c = c->geo_c; v.geo_h = c->geo_h; v.mp = c->mp;
  return v;
}

static NodeArgs node_args(gdf_ctx* c) {
  NodeArgs a; a.bbmin = c->bbmin; a.bbmax = c->bbmax; a.geo_c = c->geo_c; a.geo_h = c->geo_h;
  a.mp = c->mp;
  return a;
}

int seg_alloc(gdf_ctx* c) {
  if (c->d_seg_off) return GDF_OK;
  GDF_HIP(c, hipMalloc(&c->d_seg_off, sizeof(int)*(kMaxSeg + 1)));
  GDF_HIP(c, hipMalloc(&c->seg_root, sizeof(int)*kMaxSeg));
  GDF_HIP(c, hipMalloc(&c->seg_dt, sizeof(double)*kMaxSeg));
  GDF_HIP(c, hipMalloc(&c->seg_bnd, sizeof(double)*4*kMaxSeg));
  return GDF_OK;
}

// segment offsets to the device and the per-particle segment ids
__global__ void k_fill_seg(int n0, int n1, int sg, int* __restrict__ seg) {
  const int i = n0 + blockIdx.x*BS + threadIdx.x;
  if (i < n1) seg[i] = sg;
}

int seg_upload(gdf_ctx* c) {
  GDF_TRY(seg_alloc(c));
  if (c->nseg < 1 || c->nseg > kMaxSeg || (int) c->seg_off.size() != c->nseg + 1)
    return fail(c, GDF_ERR_INVALID, "bad segment table");
  GDF_HIP(c, hipMemcpyAsync(c->d_seg_off, c->seg_off.data(), sizeof(int)*(c->nseg + 1),
                            hipMemcpyHostToDevice, c->stream));
  for (int sg = 0; sg < c->nseg; sg++) {
    const int n0 = c->seg_off[sg], n1 = c->seg_off[sg + 1];
    if (n1 > n0) k_fill_seg<<<nblocks(n1 - n0, BS), BS, 0, c->stream>>>(n0, n1, sg, c->pi[I_SEG]);
  }
  GDF_HIP(c, hipGetLastError());
  c->tree_valid = false;
  c->topo_valid = false;
  return GDF_OK;
}

int tree_alloc(gdf_ctx* c) {
  const size_t N = (size_t) c->Ncap;
  const size_t nn = 2*N;  // >= 2N-1
  GDF_HIP(c, hipMalloc(&c->keys, sizeof(uint64_t)*N));
  GDF_HIP(c, hipMalloc(&c->keys_alt, sizeof(uint64_t)*N));
  GDF_HIP(c, hipMalloc(&c->vals, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->vals_alt, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->left, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->right, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->first, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->last, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->parent, sizeof(int)*nn));
  GDF_HIP(c, hipMalloc(&c->lnk, sizeof(int4)*N));
  GDF_HIP(c, hipMalloc(&c->visit, sizeof(int)*(N + 16)));   // also 16-byte-padded flags
  // refit frontiers: phase 1 (one RB slot per chunk), chunk counts, phase 2 (N + count)
  GDF_HIP(c, hipMalloc(&c->front, sizeof(int)*(2*N + 1024 + N/1024 + 16)));
  GDF_HIP(c, hipMalloc(&c->bbmin, sizeof(double)*3*nn));
  GDF_HIP(c, hipMalloc(&c->bbmax, sizeof(double)*3*nn));
  GDF_HIP(c, hipMalloc(&c->geo_c, sizeof(double4)*nn));
  GDF_HIP(c, hipMalloc(&c->geo_h, sizeof(double4)*nn));
  GDF_HIP(c, hipMalloc(&c->mp, sizeof(double)*10*nn));
  GDF_HIP(c, hipMalloc(&c->gstart, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->grp_first, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->grp_count, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->grp_node, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->grp_root, sizeof(int)*N));
  GDF_HIP(c, hipMalloc(&c->grp_gmac, sizeof(double)*N));
  GDF_HIP(c, hipMalloc(&c->nsel, sizeof(int)*4));
  // temporary storage of the sort and the group compaction
  size_t b1 = 0, b2 = 0;
  GDF_HIP(c, rocprim::radix_sort_pairs(nullptr, b1, c->keys, c->keys_alt, c->vals, c->vals_alt,
                                       (unsigned int) N, 0, 64, c->stream));
  b2 = sizeof(int)*((size_t) nblocks(nblocks((long long) N, 16), kSelBlock) + 16);
  c->sort_tmp_bytes = b1;
  c->sel_tmp_bytes = b2;
  GDF_HIP(c, hipMalloc(&c->sort_tmp, b1 > 0 ? b1 : 16));
  GDF_HIP(c, hipMalloc(&c->sel_tmp, b2 > 0 ? b2 : 16));
  return GDF_OK;
}

namespace {

// Order-preserving selection of the set bytes of n 0/1 byte flags (zero-padded to whole
// 16-byte words): per-block counts, an exclusive scan of the block counts by one workgroup, and
// a per-block scatter of the indices.  No memset and no library call, so the tree build's
// launches capture into a graph of kernel nodes only.
__device__ __forceinline__ int bytes_set(uint4 w) {
  // each byte is 0 or 1: the byte sum of a word is its top byte after multiplying by 0x01010101
  return (int) (((w.x*0x01010101u) >> 24) + ((w.y*0x01010101u) >> 24) +
                ((w.z*0x01010101u) >> 24) + ((w.w*0x01010101u) >> 24));
}

__global__ __launch_bounds__(kSelBlock) void k_sel_count(int nw, const uint4* __restrict__ flags,
                                                         int* __restrict__ bcount) {
  __shared__ int red[kSelBlock];
  const int w = blockIdx.x*kSelBlock + threadIdx.x;
  red[threadIdx.x] = w < nw ? bytes_set(flags[w]) : 0;
  __syncthreads();
  for (int st = kSelBlock/2; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) bcount[blockIdx.x] = red[0];
}

// exclusive scan of nb block counts in place, the total into *total; one workgroup
__global__ __launch_bounds__(kSelBlock) void k_sel_scan(int nb, int* __restrict__ bcount,
                                                        int* __restrict__ total) {
  __shared__ int part[kSelBlock];
  const int per = (nb + kSelBlock - 1)/kSelBlock;
  const int b0 = min(threadIdx.x*per, nb), b1 = min(b0 + per, nb);
  int sum = 0;
  for (int b = b0; b < b1; b++) sum += bcount[b];
  part[threadIdx.x] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int t = 0; t < kSelBlock; t++) { const int v = part[t]; part[t] = run; run += v; }
    *total = run;
  }
  __syncthreads();
  int run = part[threadIdx.x];
  for (int b = b0; b < b1; b++) { const int v = bcount[b]; bcount[b] = run; run += v; }
}

__global__ __launch_bounds__(kSelBlock) void k_sel_scatter(int nw, const uint4* __restrict__ flags,
                                                           const int* __restrict__ boff,
                                                           int* __restrict__ out) {
  __shared__ int pre[kSelBlock];
  const int w = blockIdx.x*kSelBlock + threadIdx.x;
  const uint4 x = w < nw ? flags[w] : make_uint4(0u, 0u, 0u, 0u);
  const int cnt = bytes_set(x);
  pre[threadIdx.x] = cnt;
  __syncthreads();
  for (int st = 1; st < kSelBlock; st <<= 1) {          // inclusive scan (Hillis-Steele)
    const int v = threadIdx.x >= st ? pre[threadIdx.x - st] : 0;
    __syncthreads();
    pre[threadIdx.x] += v;
    __syncthreads();
  }
  if (cnt == 0) return;
  int o = boff[blockIdx.x] + pre[threadIdx.x] - cnt;
  const unsigned int word[4] = {x.x, x.y, x.z, x.w};
  for (int k = 0; k < 16; k++)
    if ((word[k >> 2] >> (8*(k & 3))) & 0xffu) out[o++] = 16*w + k;
}

// The byte flags of the group selection, cleared by a kernel (a memset would become a memset
// node of the captured graph)
__global__ void k_zero_words(int n, uint4* __restrict__ p) {
  const int i = blockIdx.x*BS + threadIdx.x;
  if (i < n) p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// Steps 1-6 of a tree build: every launch that needs no value back on the host.  The host's
// buffer swaps (double-buffered keys, indices and particle state) run with launch == false
// too, so a replayed graph of these launches leaves the host state as the launches themselves
// would.
int build_body(gdf_ctx* c, hipStream_t s, bool launch) {
  const int N = c->N;
  const int nbp = std::min(nblocks(N, BS), 1024);
  // 1. bounds and Morton keys
  if (launch) {
    if (c->nseg > 0) {
      k_seg_bounds_partial<<<dim3(kSegBlocks, c->nseg), BS, 0, s>>>(c->d_seg_off, c->pf[F_X],
                                                                    c->pf[F_Y], c->pf[F_Z],
                                                                    c->red);
      k_seg_bounds_final<<<c->nseg, 64, 0, s>>>(c->red, c->seg_bnd);
      k_morton_seg<<<nblocks(N, BS), BS, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z],
                                                 c->pi[I_SEG], c->seg_bnd, c->keys, c->vals);
    } else {
      k_bounds_partial<<<nbp, BS, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z], c->red);
      k_bounds_final<<<1, BS, 0, s>>>(nbp, c->red, c->red + 6*1024);
      k_morton<<<nblocks(N, BS), BS, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z],
                                             c->red + 6*1024, c->keys, c->vals);
    }
    // 2. stable radix sort of (key, index)
    size_t bytes = c->sort_tmp_bytes;
    GDF_HIP(c, rocprim::radix_sort_pairs(c->sort_tmp, bytes, c->keys, c->keys_alt, c->vals,
                                         c->vals_alt, (unsigned int) N, 0,
                                         c->nseg > 0 ? 64 : 63, s));
  }
  std::swap(c->keys, c->keys_alt);
  std::swap(c->vals, c->vals_alt);
  // 3. permute persistent particle state into key order
  if (launch) {
    PermArgs pa;
    for (int f = 0; f < F_NPERSIST; f++) { pa.src[f] = c->pf[f]; pa.dst[f] = c->pf_alt[f]; }
    for (int f = 0; f < I_NINT; f++) { pa.isrc[f] = c->pi[f]; pa.idst[f] = c->pi_alt[f]; }
    k_permute<<<nblocks(N, BS), BS, 0, s>>>(N, c->vals, pa, c->iorig, c->iorig_alt);
  }
  for (int f = 0; f < F_NPERSIST; f++) std::swap(c->pf[f], c->pf_alt[f]);
  for (int f = 0; f < I_NINT; f++) std::swap(c->pi[f], c->pi_alt[f]);
  std::swap(c->iorig, c->iorig_alt);
  if (c->df_alt[0]) {
    if (launch) {
      PermDArgs pd;
      for (int k = 0; k < kNPermD; k++) {
        pd.src[k] = c->df[kPermD[k]]; pd.dst[k] = c->df_alt[kPermD[k]];
      }
      k_permute_d<<<nblocks(N, BS), BS, 0, s>>>(N, c->vals, pd);
    }
    for (int k = 0; k < kNPermD; k++) std::swap(c->df[kPermD[k]], c->df_alt[kPermD[k]]);
  }
  if (!launch) return GDF_OK;
  // 4. topology
  if (N > 1)
    k_lbvh<<<nblocks(N - 1, BS), BS, 0, s>>>(N, c->keys, c->left, c->right, c->first, c->last,
                                             c->parent, c->lnk);
  // 5. moments, bottom-up
  NodeArgs na = node_args(c);
  const double theta2 = c->phys.thetamaxsqd;
  if (N > 1)
    launch_refit<false>(c, na, theta2);
  else
    k_leaves<<<1, BS, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z], c->pf[F_M], c->pf[F_H],
                              theta2, na, c->visit, c->parent);
  // 6. walk groups in ascending first-particle order
  const int nw = (N + 15)/16;                // N byte flags, padded to 16 bytes
  k_zero_words<<<nblocks(nw, BS), BS, 0, s>>>(nw, reinterpret_cast<uint4*>(c->visit));
  k_groups_mark<<<nblocks(2*N - 1, BS), BS, 0, s>>>(N, c->cfg.group_max, c->parent, c->first,
                                                    c->last, c->gstart,
                                                    (unsigned char*) c->visit);
  const int nsb = nblocks(nw, kSelBlock);
  int* bcount = reinterpret_cast<int*>(c->sel_tmp);
  const uint4* fw = reinterpret_cast<const uint4*>(c->visit);
  k_sel_count<<<nsb, kSelBlock, 0, s>>>(nw, fw, bcount);
  k_sel_scan<<<1, kSelBlock, 0, s>>>(nsb, bcount, c->nsel);
  k_sel_scatter<<<nsb, kSelBlock, 0, s>>>(nw, fw, bcount, c->grp_first);
  k_groups_fill<<<nblocks(N, BS), BS, 0, s>>>(N, c->nsel, c->grp_first, c->gstart, c->first,
                                              c->last, c->grp_node, c->grp_count);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

// every buffer pointer the build swaps, in a fixed order: with the particle count, segment
// count and configuration of the context it keys a captured graph
std::vector<const void*> build_key(const gdf_ctx* c) {
  std::vector<const void*> k = {c->keys, c->keys_alt, c->vals, c->vals_alt, c->iorig,
                                c->iorig_alt, c->df_alt[0]};
  for (int f = 0; f < F_NPERSIST; f++) { k.push_back(c->pf[f]); k.push_back(c->pf_alt[f]); }
  for (int f = 0; f < I_NINT; f++) { k.push_back(c->pi[f]); k.push_back(c->pi_alt[f]); }
  for (int i = 0; i < kNPermD; i++) {
    k.push_back(c->df[kPermD[i]]); k.push_back(c->df_alt[kPermD[i]]);
  }
  // every other buffer and value the launches take
  const void* other[] = {c->d_seg_off, c->seg_bnd, c->red, c->sort_tmp, c->sel_tmp, c->nsel,
                         c->left, c->right, c->first, c->last, c->parent, c->lnk, c->visit, c->front,
                         c->bbmin, c->bbmax, c->geo_c, c->geo_h, c->mp, c->gstart, c->grp_first,
                         c->grp_count, c->grp_node};
  for (const void* q : other) k.push_back(q);
  uint64_t theta_bits;
  std::memcpy(&theta_bits, &c->phys.thetamaxsqd, sizeof(theta_bits));
  const uint64_t vals[] = {(uint64_t) c->N, (uint64_t) c->nseg, (uint64_t) c->sort_tmp_bytes,
                           (uint64_t) c->sel_tmp_bytes, (uint64_t) c->cfg.group_max, theta_bits};
  for (uint64_t v : vals) k.push_back(reinterpret_cast<const void*>((uintptr_t) v));
  return k;
}

// a captured graph is used only if it holds kernel and copy nodes (no memset nodes, whose
// replays are not trusted here)
// returns 0, or 100 + the type of the first node of another kind
int graph_nodes_check(hipGraph_t g) {
  size_t n = 0;
  if (hipGraphGetNodes(g, nullptr, &n) != hipSuccess) return 99;
  std::vector<hipGraphNode_t> nodes(n);
  if (n > 0 && hipGraphGetNodes(g, nodes.data(), &n) != hipSuccess) return 99;
  for (hipGraphNode_t nd : nodes) {
    hipGraphNodeType t;
    if (hipGraphNodeGetType(nd, &t) != hipSuccess) return 99;
    if (t != hipGraphNodeTypeKernel && t != hipGraphNodeTypeMemcpy &&
        t != hipGraphNodeTypeEmpty)
      return 100 + (int) t;
  }
  return 0;
}

// the build's launches through a graph captured at the first build of each buffer
// arrangement (the swaps alternate between two) and replayed after; a capture that fails or
// holds other nodes leaves the context building eagerly
int build_launches(gdf_ctx* c, hipStream_t s) {
  if (!c->cfg.graph_capture || c->graph_failed) return build_body(c, s, true);
  const std::vector<const void*> key = build_key(c);
  for (int e = 0; e < gdf_ctx::kBuildGraphs; e++)
    if (c->bg_exec[e] && c->bg_key[e] == key) {
      GDF_HIP(c, hipGraphLaunch(c->bg_exec[e], s));
      c->graph_replays++;
      return build_body(c, s, false);
    }
  // capture: the host swaps happen during the capture, so they are undone afterwards (the
  // replay below redoes them)
  struct Saved {
    double* pf[F_NPERSIST]; double* pf_alt[F_NPERSIST];
    int* pi[I_NINT]; int* pi_alt[I_NINT];
    double* df[D_NDERIVED]; double* df_alt[D_NDERIVED];
    uint64_t* keys; uint64_t* keys_alt; int* vals; int* vals_alt; int* iorig; int* iorig_alt;
  } sv;
  std::copy(std::begin(c->pf), std::end(c->pf), sv.pf);
  std::copy(std::begin(c->pf_alt), std::end(c->pf_alt), sv.pf_alt);
  std::copy(std::begin(c->pi), std::end(c->pi), sv.pi);
  std::copy(std::begin(c->pi_alt), std::end(c->pi_alt), sv.pi_alt);
  std::copy(std::begin(c->df), std::end(c->df), sv.df);
  std::copy(std::begin(c->df_alt), std::end(c->df_alt), sv.df_alt);
  sv.keys = c->keys; sv.keys_alt = c->keys_alt; sv.vals = c->vals; sv.vals_alt = c->vals_alt;
  sv.iorig = c->iorig; sv.iorig_alt = c->iorig_alt;
  auto restore = [&]() {
    std::copy(sv.pf, sv.pf + F_NPERSIST, c->pf);
    std::copy(sv.pf_alt, sv.pf_alt + F_NPERSIST, c->pf_alt);
    std::copy(sv.pi, sv.pi + I_NINT, c->pi);
    std::copy(sv.pi_alt, sv.pi_alt + I_NINT, c->pi_alt);
    std::copy(sv.df, sv.df + D_NDERIVED, c->df);
    std::copy(sv.df_alt, sv.df_alt + D_NDERIVED, c->df_alt);
    c->keys = sv.keys; c->keys_alt = sv.keys_alt; c->vals = sv.vals; c->vals_alt = sv.vals_alt;
    c->iorig = sv.iorig; c->iorig_alt = sv.iorig_alt;
  };
  hipGraph_t g = nullptr;
  hipGraphExec_t ex = nullptr;
  // why a capture was given up, kept for gdf_graph_state: 1 begin, 2 a launch, 3 end,
  // 100 + node type of a node that is not a kernel or copy, 4 instantiation
  int why = 0;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) != hipSuccess) {
    why = 1;
  } else {
    const int rc = build_body(c, s, true);
    const hipError_t ee = hipStreamEndCapture(s, &g);
    if (rc != GDF_OK) why = 2;
    else if (ee != hipSuccess || !g) why = 3;
    else if (const int nt = graph_nodes_check(g)) why = nt;
    else if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) why = 4;
  }
  const bool ok = why == 0;
  c->graph_why = why;
  if (g) (void) hipGraphDestroy(g);
  // a capture invalidated part way (by another thread's call on the device, say) may leave the
  // stream capturing: end it, and clear the error it left, before building eagerly
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
    hipGraph_t g2 = nullptr;
    (void) hipStreamEndCapture(s, &g2);
    if (g2) (void) hipGraphDestroy(g2);
  }
  (void) hipGetLastError();
  restore();
  if (!ok) {
    if (ex) (void) hipGraphExecDestroy(ex);
    c->graph_failed = true;
    return build_body(c, s, true);
  }
  const int e = c->bg_next;
  c->bg_next = (c->bg_next + 1)%gdf_ctx::kBuildGraphs;
  if (c->bg_exec[e]) (void) hipGraphExecDestroy(c->bg_exec[e]);
  c->bg_exec[e] = ex;
  c->bg_key[e] = key;
  GDF_HIP(c, hipGraphLaunch(ex, s));
  return build_body(c, s, false);
}

}  // namespace

void build_graphs_free(gdf_ctx* c) {
  for (int e = 0; e < gdf_ctx::kBuildGraphs; e++) {
    if (c->bg_exec[e]) (void) hipGraphExecDestroy(c->bg_exec[e]);
    c->bg_exec[e] = nullptr;
    c->bg_key[e].clear();
  }
}

int tree_build(gdf_ctx* c) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (N <= 0) return fail(c, GDF_ERR_INVALID, "tree build without particles");
  GDF_TRY(build_launches(c, s));
  GDF_HIP(c, hipMemcpyAsync(c->h_counters, c->nsel, sizeof(int), hipMemcpyDeviceToHost, s));
  GDF_HIP(c, hipStreamSynchronize(s));
  c->ngroups = c->h_counters[0];
  if (c->ngroups <= 0 || c->ngroups > N) return fail(c, GDF_ERR_INVALID, "walk group build failed");
  if (c->nseg > 0) {
    k_seg_roots<<<nblocks(2*N - 1, BS), BS, 0, s>>>(N, c->nseg, c->d_seg_off, c->first, c->last,
                                                    c->pi[I_SEG], c->seg_root);
    k_grp_roots<<<nblocks(c->ngroups, BS), BS, 0, s>>>(c->ngroups, c->grp_first, c->pi[I_SEG],
                                                       c->seg_root, c->grp_root);
  }
  // the force arrays follow the new order only when the build permuted them
  GDF_TRY(group_mac(c, c->df_alt[0] ? nullptr : c->vals));
  c->tree_valid = true;
  c->topo_valid = true;
  c->sg_valid = false; c->sg_union = false;
  GDF_HIP(c, hipGetLastError());
  return integ_pack_posm(c);
}

// The kept topology and walk groups with every node's box, moments and hmax recomputed from the
// particles' current state: the bottom-up pass of a build without the sort, the permutation and
// the group selection.  Synchronises the stream, as a build does.
int tree_restock(gdf_ctx* c) {
  const int N = c->N;
  hipStream_t s = c->stream;
  if (!c->topo_valid) return fail(c, GDF_ERR_INVALID, "tree refit without a built tree");
  NodeArgs na = node_args(c);
  const double theta2 = c->phys.thetamaxsqd;
  if (N > 1)
    launch_refit<false>(c, na, theta2);
  else
    k_leaves<<<1, BS, 0, s>>>(N, c->pf[F_X], c->pf[F_Y], c->pf[F_Z], c->pf[F_M], c->pf[F_H],
                              theta2, na, c->visit, c->parent);
  GDF_HIP(c, hipGetLastError());
  GDF_TRY(group_mac(c, nullptr));
  GDF_TRY(integ_pack_posm(c));
  GDF_HIP(c, hipStreamSynchronize(s));
  c->tree_valid = true;
  c->sg_valid = false; c->sg_union = false;
  return GDF_OK;
}

int group_list(gdf_ctx* c, const int* status, int* list, int* ncand, int* gnc, int* gnn,
               int* zero, int nzero) {
  const int ng = c->ngroups;
  hipStream_t s = c->stream;
  // the refit's arrival counters are free between tree passes (phase 1 clears them)
  unsigned char* flags = reinterpret_cast<unsigned char*>(c->visit);
  constexpr int gpb = BS/kWave;   // groups per block
  k_group_flags<<<(ng + gpb - 1)/gpb, BS, 0, s>>>(ng, c->grp_first, c->grp_count, status, flags,
                                                  ncand, gnc, gnn, zero, nzero);
  k_compact_flags<<<1, kWave, 0, s>>>(ng, reinterpret_cast<const uint4*>(flags), list,
                                      c->Ncap);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

int tree_refit_h(gdf_ctx* c) {
  const int N = c->N;
  hipStream_t s = c->stream;
  NodeArgs na = node_args(c);
  const double theta2 = c->phys.thetamaxsqd;
  if (N > 1)
    launch_refit<true>(c, na, theta2);
  else
    k_hleaves<<<1, BS, 0, s>>>(N, c->pf[F_H], theta2, na, c->visit);
  GDF_HIP(c, hipGetLastError());
  return GDF_OK;
}

}  // namespace gdf
