// Blocked multi-vector kernels of the batched solver.  The FP64 blocks of one
// timeslice are stored site-major with the vectors fastest:
//     X[ts][(s*3 + colour)*K + col]            (K a multiple of 64)
// (the layout the strided-batched GEMMs read as a K x n matrix), the FP32
// filter buffers block-major, one contiguous array per block of bw = 64 or
// 128 vectors (chosen by the batch's autotuner):
//     F[ts][col/bw][(s*3 + colour)*bw + col%bw]
// so one 64-lane wavefront holds 64 vectors of ONE site: the links of that
// site are wave-uniform and are fetched once per wavefront through scalar
// loads (the link matrix is read once per 64 vectors), while every vector
// access is a fully coalesced 512-byte (FP32) / 1 KiB (FP64) row segment.
// Links per timeslice are AoS [s][dir][9] so a site's 3 forward matrices are
// one contiguous 216 B (FP32) / 432 B (FP64) scalar fetch.
//
// Work-group = 4 wavefronts = 4 consecutive sites of one 64-vector block of
// one timeslice (a "unit" = one (timeslice, 64-vector block) pair).
// Workgroups are handed to the 8 XCDs round-robin by linear id; the 1D grid
// is decoded so that every XCD owns whole units and sweeps each unit plane by
// plane (site order x-major).  A vector is then reused as the x-, y- and
// z-neighbour of later sites while it is still in that XCD's L2: the reuse
// distance is two x-planes (2 * L*L sites * 1.5 KiB).  The streamed arrays
// (r, f_{k-1}, the output) are accessed non-temporally so they do not evict
// those planes.
#include <hip/hip_runtime.h>

#include "kernels_block.h"

namespace laph {

// Decodes blockIdx.x of a 1D grid of SB site-blocks x nb column blocks x nts
// timeslices into (ts, cb, sb), giving each XCD (linear id mod 8) whole units.
__device__ __forceinline__ void unit_map(int SB, int nb, int nts, int& ts, int& cb, int& sb) {
  const unsigned id = blockIdx.x;
  const unsigned U = (unsigned)nb * nts;
  unsigned unit;
  if ((U & 7u) == 0) {
    const unsigned j = id >> 3;
    sb = (int)(j % SB);
    unit = (j / SB) * 8 + (id & 7u);
  } else {
    sb = (int)(id % SB);
    unit = id / SB;
  }
  ts = (int)(unit / nb);
  cb = (int)(unit % nb);
}

// Site of sweep index i when each unit is swept in y-strips of width W:
// i = ((strip * L + x) * W + yy) * L + z, y = strip * W + yy (W divides L;
// W = L is plain x-major order).  With narrow strips the x-neighbour planes
// of the sites in flight on an XCD (W * L sites each) stay in its L2.
__device__ __forceinline__ int strip_site(int i, int L, int W) {
  const int LW = L * W;
  const int strip = i / (LW * L);
  const int r = i - strip * (LW * L);
  const int x = r / LW;
  const int r2 = r - x * LW;
  const int yy = r2 / L;
  const int z = r2 - yy * L;
  return (x * L + strip * W + yy) * L + z;
}

typedef float f2v __attribute__((ext_vector_type(2)));
typedef double d2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld_nt(const float2* p) {
  const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(p));
  return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st_nt(float2* p, float2 v) {
  f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<f2v*>(p));
}
typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld_nt4(const float4* p) {
  const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(p));
  return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void st_nt4(float4* p, float4 v) {
  f4v w;
  w.x = v.x;
  w.y = v.y;
  w.z = v.z;
  w.w = v.w;
  __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ void st_nt(double2* p, double2 v) {
  d2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, reinterpret_cast<d2v*>(p));
}

template <class C>
__device__ __forceinline__ C cfma(C u, C v, C acc) {  // acc += u*v
  acc.x = fma(u.x, v.x, fma(-u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(u.y, v.x, acc.y));
  return acc;
}
template <class C>
__device__ __forceinline__ C cfmac(C u, C v, C acc) {  // acc += conj(u)*v
  acc.x = fma(u.x, v.x, fma(u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(-u.y, v.x, acc.y));
  return acc;
}

// Hopping term h(s) = sum_mu [U_mu(s) v(s+mu) + U_mu(s-mu)^H v(s-mu)] for the
// lane's vector; returns h and v(s) in acc/ctr.
template <class C>
__device__ __forceinline__ void hop(const C* __restrict__ U, const C* __restrict__ v, int s, int L,
                                    int K, int lane, C acc[3], C ctr[3]) {
  const int L2 = L * L;
  const int x = s / L2, y = (s / L) % L, z = s - (s / L) * L;
  const int up[3] = {x == L - 1 ? s - (L - 1) * L2 : s + L2, y == L - 1 ? s - (L - 1) * L : s + L,
                     z == L - 1 ? s - (L - 1) : s + 1};
  const int dn[3] = {x == 0 ? s + (L - 1) * L2 : s - L2, y == 0 ? s + (L - 1) * L : s - L,
                     z == 0 ? s + (L - 1) : s - 1};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    acc[c] = C{0, 0};
    ctr[c] = v[(size_t)(3 * s + c) * K + lane];
  }
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    C xu[3], xm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xu[c] = v[(size_t)(3 * up[mu] + c) * K + lane];
      xm[c] = v[(size_t)(3 * dn[mu] + c) * K + lane];
    }
    const C* Uf = U + ((size_t)s * 3 + mu) * 9;       // wave-uniform
    const C* Ub = U + ((size_t)dn[mu] * 3 + mu) * 9;  // wave-uniform
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        acc[a] = cfma(Uf[3 * a + b], xu[b], acc[a]);
        acc[a] = cfmac(Ub[3 * b + a], xm[b], acc[a]);
      }
    }
  }
}

// One Chebyshev step in FP32 for all timeslices:
//   out = a1 * (B in) + a3 * in + a2 * r + b * old,   B = c_t (H - 6) + a_t
// with per-column coefficients (a1, a2, b, a3) for this step, evaluated as
//   out = a1 * (H' in) + w * in + a2 * r + b * old
// from links prescaled by c_t (H' = c_t H: the filter's link copy is
// rewritten each cycle) and w = a1 (a_t - 6 c_t) + a3 folded into the
// coefficient table, so each output component costs one multiply and two
// FMAs after the hopping sum.  HAS_R = false
// is the plain recurrence (no source term, one vector stream less), and a
// null old the product-form step (no previous vector), whose shift term is
// complex: a3 + i b.
template <bool HAS_R>
__global__ __launch_bounds__(256) void k_filter_step32(const float2* __restrict__ U32,
                                                        const float2* __restrict__ in,
                                                        const float2* __restrict__ r,
                                                        const float2* __restrict__ old,
                                                        float2* __restrict__ out,
                                                        const float4* __restrict__ coef,
                                                        const float2* __restrict__ cab,
                                                        const int* __restrict__ first_blk, int L,
                                                        int K, int nts, int use_old, int W) {
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 3) / 4, K / 64, nts, ts, cb, sb);
  if (cb < first_blk[ts]) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int si = sb * 4 + wave;
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int l64 = threadIdx.x & 63;
  const int lane = cb * 64 + l64;
  // FP32 filter buffers are block-major: [ts][cb][3*V3][64]
  const size_t vo = ((size_t)ts * (K / 64) + cb) * (size_t)V3 * 3 * 64;
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  float2 acc[3], ctr[3];
  hop<float2>(Ut, in + vo, s, L, 64, l64, acc, ctr);
  const float4 cf = coef[(size_t)ts * K + lane];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t idx = vo + (size_t)(3 * s + c) * 64 + l64;
    float2 o;
    o.x = cf.x * acc[c].x + cf.w * ctr[c].x;
    o.y = cf.x * acc[c].y + cf.w * ctr[c].y;
    if (HAS_R) {
      const float2 rv = ld_nt(r + idx);
      o.x += cf.y * rv.x;
      o.y += cf.y * rv.y;
    }
    if (use_old) {
      const float2 ov = ld_nt(old + idx);
      o.x += cf.z * ov.x;
      o.y += cf.z * ov.y;
    } else {  // imaginary part of a complex shift
      o.x -= cf.z * ctr[c].y;
      o.y += cf.z * ctr[c].x;
    }
    st_nt(out + idx, o);
  }
}

// 32-vector-block variant (filter buffers block-major with 32-vector blocks,
// 256 B per site row): each wavefront holds two sites, one per half, so a
// work-group covers 8 consecutive sites of the sweep.  The two halves read
// different sites' links, through vector loads instead of wave-uniform scalar
// loads; the links are read once per 32 vectors.
template <bool HAS_R>
__global__ __launch_bounds__(256) void k_filter_step32h(const float2* __restrict__ U32,
                                                         const float2* __restrict__ in,
                                                         const float2* __restrict__ r,
                                                         const float2* __restrict__ old,
                                                         float2* __restrict__ out,
                                                         const float4* __restrict__ coef,
                                                         const float2* __restrict__ cab,
                                                         const int* __restrict__ first_blk, int L,
                                                         int K, int nts, int use_old, int W) {
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 7) / 8, K / 32, nts, ts, cb, sb);
  if (cb * 32 + 31 < first_blk[ts] * 64) return;  // the whole 32-vector block is locked
  const int l32 = threadIdx.x & 31;
  const int si = sb * 8 + (threadIdx.x >> 5);
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int col = cb * 32 + l32;
  const size_t vo = ((size_t)ts * (K / 32) + cb) * (size_t)V3 * 3 * 32;
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  float2 acc[3], ctr[3];
  hop<float2>(Ut, in + vo, s, L, 32, l32, acc, ctr);
  const float4 cf = coef[(size_t)ts * K + col];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t idx = vo + (size_t)(3 * s + c) * 32 + l32;
    float2 o;
    o.x = cf.x * acc[c].x + cf.w * ctr[c].x;
    o.y = cf.x * acc[c].y + cf.w * ctr[c].y;
    if (HAS_R) {
      const float2 rv = ld_nt(r + idx);
      o.x += cf.y * rv.x;
      o.y += cf.y * rv.y;
    }
    if (use_old) {
      const float2 ov = ld_nt(old + idx);
      o.x += cf.z * ov.x;
      o.y += cf.z * ov.y;
    } else {  // imaginary part of a complex shift
      o.x -= cf.z * ctr[c].y;
      o.y += cf.z * ctr[c].x;
    }
    st_nt(out + idx, o);
  }
}

// 128-vector-block variant: each lane carries two neighbouring vectors as one
// 16-byte float4 (buffers block-major with 128-vector blocks), halving the
// wavefronts and the load instructions per vector.
template <bool HAS_R>
__global__ __launch_bounds__(256) void k_filter_step32x2(const float2* __restrict__ U32,
                                                          const float4* __restrict__ in,
                                                          const float4* __restrict__ r,
                                                          const float4* __restrict__ old,
                                                          float4* __restrict__ out,
                                                          const float4* __restrict__ coef,
                                                          const float2* __restrict__ cab,
                                                          const int* __restrict__ first_blk, int L,
                                                          int K, int nts, int use_old, int W) {
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 3) / 4, K / 128, nts, ts, cb, sb);
  if (2 * cb + 1 < first_blk[ts]) return;  // both 64-vector halves inactive
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int si = sb * 4 + wave;
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int l64 = threadIdx.x & 63;
  const int col = cb * 128 + 2 * l64;
  if (col < first_blk[ts] * 64) return;  // lanes of a locked 64-vector half issue no traffic
  const size_t vo = ((size_t)ts * (K / 128) + cb) * (size_t)V3 * 3 * 64;  // in float4 units
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int x = s / L2, y = (s / L) % L, z = s - (s / L) * L;
  const int up[3] = {x == L - 1 ? s - (L - 1) * L2 : s + L2, y == L - 1 ? s - (L - 1) * L : s + L,
                     z == L - 1 ? s - (L - 1) : s + 1};
  const int dn[3] = {x == 0 ? s + (L - 1) * L2 : s - L2, y == 0 ? s + (L - 1) * L : s - L,
                     z == 0 ? s + (L - 1) : s - 1};
  const float4* v = in + vo;
  float2 acc0[3], acc1[3];
  float4 ctr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    acc0[c] = make_float2(0.f, 0.f);
    acc1[c] = make_float2(0.f, 0.f);
    ctr[c] = v[(size_t)(3 * s + c) * 64 + l64];
  }
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    float4 xu[3], xm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xu[c] = v[(size_t)(3 * up[mu] + c) * 64 + l64];
      xm[c] = v[(size_t)(3 * dn[mu] + c) * 64 + l64];
    }
    const float2* Uf = Ut + ((size_t)s * 3 + mu) * 9;
    const float2* Ub = Ut + ((size_t)dn[mu] * 3 + mu) * 9;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const float2 uf = Uf[3 * a + bb], ub = Ub[3 * bb + a];
        acc0[a] = cfma(uf, make_float2(xu[bb].x, xu[bb].y), acc0[a]);
        acc1[a] = cfma(uf, make_float2(xu[bb].z, xu[bb].w), acc1[a]);
        acc0[a] = cfmac(ub, make_float2(xm[bb].x, xm[bb].y), acc0[a]);
        acc1[a] = cfmac(ub, make_float2(xm[bb].z, xm[bb].w), acc1[a]);
      }
    }
  }
  const float4 cf0 = coef[(size_t)ts * K + col], cf1 = coef[(size_t)ts * K + col + 1];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t idx = vo + (size_t)(3 * s + c) * 64 + l64;
    float4 o;
    o.x = cf0.x * acc0[c].x + cf0.w * ctr[c].x;
    o.y = cf0.x * acc0[c].y + cf0.w * ctr[c].y;
    o.z = cf1.x * acc1[c].x + cf1.w * ctr[c].z;
    o.w = cf1.x * acc1[c].y + cf1.w * ctr[c].w;
    if (HAS_R) {
      const float4 rv = ld_nt4(r + idx);
      o.x += cf0.y * rv.x;
      o.y += cf0.y * rv.y;
      o.z += cf1.y * rv.z;
      o.w += cf1.y * rv.w;
    }
    if (use_old) {
      const float4 ov = ld_nt4(old + idx);
      o.x += cf0.z * ov.x;
      o.y += cf0.z * ov.y;
      o.z += cf1.z * ov.z;
      o.w += cf1.z * ov.w;
    } else {  // imaginary part of a complex shift
      o.x -= cf0.z * ctr[c].y;
      o.y += cf0.z * ctr[c].x;
      o.z -= cf1.z * ctr[c].w;
      o.w += cf1.z * ctr[c].z;
    }
    st_nt4(out + idx, o);
  }
}

// Link terms of one direction for the lane's two vectors: acc += U_f xu + U_b^H xm.
__device__ __forceinline__ void link_terms2(const float2* __restrict__ Uf, const float2* __restrict__ Ub,
                                            const float4 xu[3], const float4 xm[3], float2 acc0[3],
                                            float2 acc1[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      const float2 uf = Uf[3 * a + bb], ub = Ub[3 * bb + a];
      acc0[a] = cfma(uf, make_float2(xu[bb].x, xu[bb].y), acc0[a]);
      acc1[a] = cfma(uf, make_float2(xu[bb].z, xu[bb].w), acc1[a]);
      acc0[a] = cfmac(ub, make_float2(xm[bb].x, xm[bb].y), acc0[a]);
      acc1[a] = cfmac(ub, make_float2(xm[bb].z, xm[bb].w), acc1[a]);
    }
  }
}

// Row c of site `site` of a block-major 128-vector buffer (3 rows of 64
// float4 per site): a wave-uniform pointer when site is, so the address is
// formed in scalar registers and the lane index is a 32-bit offset.
__device__ __forceinline__ const float4* row(const float4* p, int site, int c) {
  return p + (unsigned)((3 * site + c) * 64);
}

// The 128-vector-block step with a work-group of 4 wavefronts on 4
// neighbouring y-rows of one x-plane of a sweep strip (the strip width W is a
// multiple of 4), each wavefront walking the whole z-line of its row in
// unrolled runs of NZ sites (NZ divides L).  The z-1, z and z+1 rows of the
// lane's vectors are kept in registers and shifted along the line; each site's
// centre rows go through LDS, so the y-neighbours of the inner rows come from
// the neighbouring wavefronts and only the outer rows of the group load them:
// 10.5 instead of 21 global row loads per site.  The work-groups of one unit
// run y-block, then x, then strip, so the lines in flight on an XCD are the
// neighbouring x-planes of one strip, all advancing along z.  One work-group
// barrier per site (LDS double-buffered by site parity: a buffer is rewritten
// two sites later, after every wavefront has passed the barrier between).
// Lanes of a locked 64-vector half
// (only in the one block per timeslice that straddles the lock boundary) load
// and compute like the others and only skip their stores, so the wavefronts
// never diverge.  Site indices are wave-uniform: row addresses and links are
// formed in scalar registers.
template <bool HAS_R, bool OLD, int NZ>
__global__ __launch_bounds__(256) void k_filter_step32x2y(const float2* __restrict__ U32,
                                                           const float4* __restrict__ in,
                                                           const float4* __restrict__ r,
                                                           const float4* __restrict__ old,
                                                           float4* __restrict__ out,
                                                           const float4* __restrict__ coef,
                                                           const float2* __restrict__ cab,
                                                           const int* __restrict__ first_blk, int L,
                                                           int K, int nts, int W) {
  __shared__ float4 rows[2][4][3][64];
  const int V3 = L * L * L;
  const int nyb = W / 4;
  int ts, cb, g;
  unit_map(V3 / (4 * L), K / 128, nts, ts, cb, g);
  const int fbl = first_blk[ts];
  if (2 * cb + 1 < fbl) return;  // whole work-group: both 64-vector halves inactive
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int yb = g % nyb;
  const int t = g / nyb;
  const int x = t % L, strip = t / L;
  const int y = strip * W + yb * 4 + wave;
  const int l64 = threadIdx.x & 63;
  const int col = cb * 128 + 2 * l64;
  const bool act = col >= fbl * 64;
  const size_t vo = ((size_t)ts * (K / 128) + cb) * (size_t)V3 * 3 * 64;  // in float4 units
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int line = (x * L + y) * L;  // site z = 0 of the wavefront's line
  const int xup = x == L - 1 ? line - (L - 1) * L2 : line + L2;
  const int xdn = x == 0 ? line + (L - 1) * L2 : line - L2;
  const int yup = y == L - 1 ? line - (L - 1) * L : line + L;
  const int ydn = y == 0 ? line + (L - 1) * L : line - L;
  const float4* v = in + vo;
  const float4 cf0 = coef[(size_t)ts * K + col], cf1 = coef[(size_t)ts * K + col + 1];
  float4 zm[3], zc[3];
  {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      zm[c] = row(v, line + L - 1, c)[l64];
      zc[c] = row(v, line, c)[l64];
    }
  }
  for (int z0 = 0; z0 < L; z0 += NZ) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int z = z0 + k;
      const int par = z & 1;
      const int s = line + z;
      const int zdn = line + (z == 0 ? L - 1 : z - 1);
      const int zup = line + (z == L - 1 ? 0 : z + 1);
      float4 zn[3], yu[3], ym[3], xu[3], xm[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) rows[par][wave][c][l64] = zc[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zn[c] = row(v, zup, c)[l64];
        xu[c] = row(v, xup + z, c)[l64];
        xm[c] = row(v, xdn + z, c)[l64];
      }
      if (wave == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yu[c] = row(v, yup + z, c)[l64];
      }
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ym[c] = row(v, ydn + z, c)[l64];
      }
      // x and z terms need no other wavefront's rows: computed before the
      // barrier, so a wavefront waits there with less work left after it
      float2 acc0[3], acc1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc0[c] = make_float2(0.f, 0.f);
        acc1[c] = make_float2(0.f, 0.f);
      }
      const int su = __builtin_amdgcn_readfirstlane(s), xd = __builtin_amdgcn_readfirstlane(xdn + z);
      const int yd = __builtin_amdgcn_readfirstlane(ydn + z), zd = __builtin_amdgcn_readfirstlane(zdn);
      link_terms2(Ut + ((size_t)su * 3 + 0) * 9, Ut + ((size_t)xd * 3 + 0) * 9, xu, xm, acc0, acc1);
      link_terms2(Ut + ((size_t)su * 3 + 2) * 9, Ut + ((size_t)zd * 3 + 2) * 9, zn, zm, acc0, acc1);
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (wave < 3) yu[c] = rows[par][wave + 1][c][l64];
        if (wave > 0) ym[c] = rows[par][wave - 1][c][l64];
      }
      link_terms2(Ut + ((size_t)su * 3 + 1) * 9, Ut + ((size_t)yd * 3 + 1) * 9, yu, ym, acc0, acc1);
      float4 o[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const unsigned ro = (unsigned)((3 * su + c) * 64);
        const float4 ctr = zc[c];
        o[c].x = cf0.x * acc0[c].x + cf0.w * ctr.x;
        o[c].y = cf0.x * acc0[c].y + cf0.w * ctr.y;
        o[c].z = cf1.x * acc1[c].x + cf1.w * ctr.z;
        o[c].w = cf1.x * acc1[c].y + cf1.w * ctr.w;
        if (HAS_R) {
          const float4 rv = ld_nt4(r + vo + ro + l64);
          o[c].x += cf0.y * rv.x;
          o[c].y += cf0.y * rv.y;
          o[c].z += cf1.y * rv.z;
          o[c].w += cf1.y * rv.w;
        }
        if (OLD) {
          const float4 ov = ld_nt4(old + vo + ro + l64);
          o[c].x += cf0.z * ov.x;
          o[c].y += cf0.z * ov.y;
          o[c].z += cf1.z * ov.z;
          o[c].w += cf1.z * ov.w;
        } else {  // imaginary part of a complex shift
          o[c].x -= cf0.z * ctr.y;
          o[c].y += cf0.z * ctr.x;
          o[c].z -= cf1.z * ctr.w;
          o[c].w += cf1.z * ctr.z;
        }
      }
      if (act) {
#pragma unroll
        for (int c = 0; c < 3; ++c) st_nt4(out + vo + (unsigned)((3 * su + c) * 64) + l64, o[c]);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zm[c] = zc[c];
        zc[c] = zn[c];
      }
    }
  }
}

// One vector per lane: link products of one 64-vector block.
__device__ __forceinline__ void link_terms1(const float2* __restrict__ Uf, const float2* __restrict__ Ub,
                                            const float2 xu[3], const float2 xm[3], float2 acc[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
#pragma unroll
    for (int bb = 0; bb < 3; ++bb) {
      acc[a] = cfma(Uf[3 * a + bb], xu[bb], acc[a]);
      acc[a] = cfmac(Ub[3 * bb + a], xm[bb], acc[a]);
    }
  }
}

// The y-row work-group step of k_filter_step32x2y on 64-vector blocks: one
// vector per lane (float2 rows of 512 B), so a wavefront holds half the
// registers and twice as many wavefronts fit on a SIMD; the links are read
// once per 64 vectors instead of once per 128.  Same walk, LDS exchange and
// barrier structure; work-groups of locked blocks return at once.
template <bool HAS_R, bool OLD, int NZ>
__global__ __launch_bounds__(256) void k_filter_step32y(const float2* __restrict__ U32,
                                                        const float2* __restrict__ in,
                                                        const float2* __restrict__ r,
                                                        const float2* __restrict__ old,
                                                        float2* __restrict__ out,
                                                        const float4* __restrict__ coef,
                                                        const int* __restrict__ first_blk, int L, int K, int nts,
                                                        int W) {
  __shared__ float2 rows[2][4][3][64];
  const int V3 = L * L * L;
  const int nyb = W / 4;
  int ts, cb, g;
  unit_map(V3 / (4 * L), K / 64, nts, ts, cb, g);
  if (cb < first_blk[ts]) return;  // whole work-group: locked block
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int yb = g % nyb;
  const int t = g / nyb;
  const int x = t % L, strip = t / L;
  const int y = strip * W + yb * 4 + wave;
  const int l64 = threadIdx.x & 63;
  const int col = cb * 64 + l64;
  const size_t vo = ((size_t)ts * (K / 64) + cb) * (size_t)V3 * 3 * 64;  // in float2 units
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int line = (x * L + y) * L;
  const int xup = x == L - 1 ? line - (L - 1) * L2 : line + L2;
  const int xdn = x == 0 ? line + (L - 1) * L2 : line - L2;
  const int yup = y == L - 1 ? line - (L - 1) * L : line + L;
  const int ydn = y == 0 ? line + (L - 1) * L : line - L;
  const float2* v = in + vo;
  auto rw = [&](int site, int c) { return v + (unsigned)((3 * site + c) * 64); };
  const float4 cf = coef[(size_t)ts * K + col];
  float2 zm[3], zc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zm[c] = rw(line + L - 1, c)[l64];
    zc[c] = rw(line, c)[l64];
  }
  for (int z0 = 0; z0 < L; z0 += NZ) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int z = z0 + k;
      const int par = z & 1;
      const int s = line + z;
      const int zdn = line + (z == 0 ? L - 1 : z - 1);
      const int zup = line + (z == L - 1 ? 0 : z + 1);
      float2 zn[3], yu[3], ym[3], xu[3], xm[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) rows[par][wave][c][l64] = zc[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zn[c] = rw(zup, c)[l64];
        xu[c] = rw(xup + z, c)[l64];
        xm[c] = rw(xdn + z, c)[l64];
      }
      if (wave == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yu[c] = rw(yup + z, c)[l64];
      }
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ym[c] = rw(ydn + z, c)[l64];
      }
      float2 acc[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] = make_float2(0.f, 0.f);
      const int su = __builtin_amdgcn_readfirstlane(s), xd = __builtin_amdgcn_readfirstlane(xdn + z);
      const int yd = __builtin_amdgcn_readfirstlane(ydn + z), zd = __builtin_amdgcn_readfirstlane(zdn);
      link_terms1(Ut + ((size_t)su * 3 + 0) * 9, Ut + ((size_t)xd * 3 + 0) * 9, xu, xm, acc);
      link_terms1(Ut + ((size_t)su * 3 + 2) * 9, Ut + ((size_t)zd * 3 + 2) * 9, zn, zm, acc);
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (wave < 3) yu[c] = rows[par][wave + 1][c][l64];
        if (wave > 0) ym[c] = rows[par][wave - 1][c][l64];
      }
      link_terms1(Ut + ((size_t)su * 3 + 1) * 9, Ut + ((size_t)yd * 3 + 1) * 9, yu, ym, acc);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const unsigned ro = (unsigned)((3 * su + c) * 64);
        const float2 ctr = zc[c];
        float2 o;
        o.x = cf.x * acc[c].x + cf.w * ctr.x;
        o.y = cf.x * acc[c].y + cf.w * ctr.y;
        if (HAS_R) {
          const float2 rv = ld_nt(r + vo + ro + l64);
          o.x += cf.y * rv.x;
          o.y += cf.y * rv.y;
        }
        if (OLD) {
          const float2 ov = ld_nt(old + vo + ro + l64);
          o.x += cf.z * ov.x;
          o.y += cf.z * ov.y;
        } else {  // imaginary part of a complex shift
          o.x -= cf.z * ctr.y;
          o.y += cf.z * ctr.x;
        }
        st_nt(out + vo + ro + l64, o);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zm[c] = zc[c];
        zc[c] = zn[c];
      }
    }
  }
}

// BF16 storage of the filter vectors: a complex value is two bfloat16 in one
// 32-bit word (real part in the low half); all arithmetic stays FP32.
__device__ __forceinline__ float2 bf2_unpack(unsigned u) {
  return make_float2(__uint_as_float(u << 16), __uint_as_float(u & 0xffff0000u));
}
__device__ __forceinline__ unsigned bf_round(float f) {  // to nearest, ties to even
  const unsigned a = __float_as_uint(f);
  return (a + 0x7fffu + ((a >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ unsigned bf2_pack(float2 v) { return bf_round(v.x) | (bf_round(v.y) << 16); }

template <int VPL>
struct UVec;
template <>
struct UVec<2> {
  typedef unsigned T __attribute__((ext_vector_type(2)));
};
template <>
struct UVec<4> {
  typedef unsigned T __attribute__((ext_vector_type(4)));
};

// The filter step on BF16-stored vectors: each lane carries VPL neighbouring
// vectors of a BW = 64*VPL block in one 8- or 16-byte word.
template <bool HAS_R, int VPL>
__global__ __launch_bounds__(256) void k_filter_step_bf16(const float2* __restrict__ U32,
                                                           const typename UVec<VPL>::T* __restrict__ in,
                                                           const typename UVec<VPL>::T* __restrict__ r,
                                                           const typename UVec<VPL>::T* __restrict__ old,
                                                           typename UVec<VPL>::T* __restrict__ out,
                                                           const float4* __restrict__ coef,
                                                           const float2* __restrict__ cab,
                                                           const int* __restrict__ first_blk, int L, int K,
                                                           int nts, int use_old, int W) {
  typedef typename UVec<VPL>::T T;
  constexpr int BW = 64 * VPL;
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 3) / 4, K / BW, nts, ts, cb, sb);
  if ((cb + 1) * VPL <= first_blk[ts]) return;  // every 64-vector part of the block locked
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int si = sb * 4 + wave;
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int l64 = threadIdx.x & 63;
  const int col = cb * BW + VPL * l64;
  if (col < first_blk[ts] * 64) return;  // lanes of locked 64-vector parts issue no traffic
  const size_t vo = ((size_t)ts * (K / BW) + cb) * (size_t)V3 * 3 * 64;  // in words of T
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int x = s / L2, y = (s / L) % L, z = s - (s / L) * L;
  const int up[3] = {x == L - 1 ? s - (L - 1) * L2 : s + L2, y == L - 1 ? s - (L - 1) * L : s + L,
                     z == L - 1 ? s - (L - 1) : s + 1};
  const int dn[3] = {x == 0 ? s + (L - 1) * L2 : s - L2, y == 0 ? s + (L - 1) * L : s - L,
                     z == 0 ? s + (L - 1) : s - 1};
  const T* v = in + vo;
  float2 acc[3][VPL];
  T ctr[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    ctr[c] = v[(size_t)(3 * s + c) * 64 + l64];
#pragma unroll
    for (int j = 0; j < VPL; ++j) acc[c][j] = make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    T xu[3], xm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      xu[c] = v[(size_t)(3 * up[mu] + c) * 64 + l64];
      xm[c] = v[(size_t)(3 * dn[mu] + c) * 64 + l64];
    }
    const float2* Uf = Ut + ((size_t)s * 3 + mu) * 9;
    const float2* Ub = Ut + ((size_t)dn[mu] * 3 + mu) * 9;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
      for (int bb = 0; bb < 3; ++bb) {
        const float2 uf = Uf[3 * a + bb], ub = Ub[3 * bb + a];
#pragma unroll
        for (int j = 0; j < VPL; ++j) {
          acc[a][j] = cfma(uf, bf2_unpack(xu[bb][j]), acc[a][j]);
          acc[a][j] = cfmac(ub, bf2_unpack(xm[bb][j]), acc[a][j]);
        }
      }
    }
  }
  float4 cf[VPL];
#pragma unroll
  for (int j = 0; j < VPL; ++j) cf[j] = coef[(size_t)ts * K + col + j];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t idx = vo + (size_t)(3 * s + c) * 64 + l64;
    T rv, ov, o;
    if (HAS_R) rv = __builtin_nontemporal_load(r + idx);
    if (use_old) ov = __builtin_nontemporal_load(old + idx);
#pragma unroll
    for (int j = 0; j < VPL; ++j) {
      const float2 xc = bf2_unpack(ctr[c][j]);
      float2 val = make_float2(cf[j].x * acc[c][j].x + cf[j].w * xc.x, cf[j].x * acc[c][j].y + cf[j].w * xc.y);
      if (HAS_R) {
        const float2 rr = bf2_unpack(rv[j]);
        val.x += cf[j].y * rr.x;
        val.y += cf[j].y * rr.y;
      }
      if (use_old) {
        const float2 oo = bf2_unpack(ov[j]);
        val.x += cf[j].z * oo.x;
        val.y += cf[j].z * oo.y;
      } else {  // imaginary part of a complex shift
        val.x -= cf[j].z * xc.y;
        val.y += cf[j].z * xc.x;
      }
      o[j] = bf2_pack(val);
    }
    __builtin_nontemporal_store(o, out + idx);
  }
}

template <int VPL>
static void launch_bf16(const float2* U32, const void* in, const void* r, const void* old, void* out,
                        const float4* coef, const float2* cab, const int* first_blk, int L, int K, int nts,
                        int W, hipStream_t st) {
  typedef typename UVec<VPL>::T T;
  const int V3 = L * L * L;
  dim3 g((unsigned)((V3 + 3) / 4) * (unsigned)(K / (64 * VPL)) * (unsigned)nts);
  const T* i4 = static_cast<const T*>(in);
  const T* r4 = static_cast<const T*>(r);
  const T* o4 = static_cast<const T*>(old);
  T* out4 = static_cast<T*>(out);
  if (r)
    k_filter_step_bf16<true, VPL><<<g, 256, 0, st>>>(U32, i4, r4, o4, out4, coef, cab, first_blk, L, K, nts,
                                                     old != nullptr, W);
  else
    k_filter_step_bf16<false, VPL><<<g, 256, 0, st>>>(U32, i4, r4, o4, out4, coef, cab, first_blk, L, K, nts,
                                                      old != nullptr, W);
}

// FP64 Laplacian on a block: out = H v - 6 v (= D v) for all columns.
__global__ __launch_bounds__(256) void k_lap_block64(const double2* __restrict__ U64,
                                                      const double2* __restrict__ v,
                                                      double2* __restrict__ out, int L, int K, int nts,
                                                      const int* __restrict__ first_blk, int W) {
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 3) / 4, K / 64, nts, ts, cb, sb);
  if (first_blk && cb < first_blk[ts]) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int si = sb * 4 + wave;
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int lane = cb * 64 + (threadIdx.x & 63);
  const size_t vo = (size_t)ts * V3 * 3 * K;
  double2 acc[3], ctr[3];
  hop<double2>(U64 + (size_t)ts * V3 * 27, v + vo, s, L, K, lane, acc, ctr);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double2 o;
    o.x = acc[c].x - 6.0 * ctr[c].x;
    o.y = acc[c].y - 6.0 * ctr[c].y;
    st_nt(out + vo + (size_t)(3 * s + c) * K + lane, o);
  }
}

// c_t = 2 / (lambda_l - lc_t) of a timeslice, lc_t = min(largest Ritz value
// of -D, wcap): the scale of B = c_t D + a_t in the current window (wcap is
// 0.75 lambda_l, lower for the warm-up cycle of a Rayleigh-Ritz start).
__device__ __forceinline__ double window_scale(const double* __restrict__ mu_ts, int K, double lamL, double wcap) {
  return 2.0 / (lamL - fmin(mu_ts[K - 1], wcap));
}

// Residuals of the Ritz pairs straight from the FP64 stencil: R = D X + X
// diag(mu) (mu the Ritz values of -D), stored as r = c_t R in the filter's
// block-major storage [ts][col/bw][row][bw] (FP32, or BF16 pairs), for the
// 64-column blocks >= first_blk[ts].  D X itself is not kept.
__global__ __launch_bounds__(256) void k_lap_residual64(const double2* __restrict__ U64,
                                                         const double2* __restrict__ v,
                                                         const double* __restrict__ mu, double lamL, double wcap,
                                                         void* __restrict__ rbuf, int L, int K, int nts,
                                                         const int* __restrict__ first_blk, int W, int bw,
                                                         int bf16) {
  const int V3 = L * L * L;
  int ts, cb, sb;
  unit_map((V3 + 3) / 4, K / 64, nts, ts, cb, sb);
  if (cb < first_blk[ts]) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int si = sb * 4 + wave;
  if (si >= V3) return;
  const int s = strip_site(si, L, W);
  const int col = cb * 64 + (threadIdx.x & 63);
  const size_t vo = (size_t)ts * V3 * 3 * K;
  double2 acc[3], ctr[3];
  hop<double2>(U64 + (size_t)ts * V3 * 27, v + vo, s, L, K, col, acc, ctr);
  const double* m = mu + (size_t)ts * K;
  const double c = window_scale(m, K, lamL, wcap), mc = m[col];
  const size_t n = (size_t)3 * V3;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float2 rv = make_float2((float)(c * (acc[a].x + (mc - 6.0) * ctr[a].x)),
                                  (float)(c * (acc[a].y + (mc - 6.0) * ctr[a].y)));
    const size_t ri = (((size_t)ts * (K / bw) + col / bw) * n + 3 * (size_t)s + a) * bw + col % bw;
    if (bf16)
      static_cast<unsigned*>(rbuf)[ri] = bf2_pack(rv);
    else
      static_cast<float2*>(rbuf)[ri] = rv;
  }
}

// The FP64 stencils on 64-column blocks in work-groups of 4 neighbouring
// y-rows of one x-plane of a sweep strip (W a multiple of 4), each wavefront
// walking zspan consecutive z-sites of its row (a multiple of NZ dividing L)
// in unrolled runs of NZ sites, as the filter step k_filter_step32x2y:
// z-neighbour rows shifted through registers,
// y-neighbours of the inner rows read from the neighbouring wavefronts' LDS
// copy of their centre rows (one barrier per site, LDS double-buffered by
// site parity), row addresses formed in scalar registers.  RES = false:
// out = D v in the site-major layout; RES = true: the Ritz residuals
// r = c_t (D v + v diag(mu)) in the filter storage (BW-vector blocks, FP32
// or BF16), as k_lap_residual64.
template <bool RES, int BW, bool BF16, int NZ, bool LINE>
__global__ __launch_bounds__(256) void k_lap64y(const double2* __restrict__ U64, const double2* __restrict__ v,
                                                double2* __restrict__ out, const double* __restrict__ mu,
                                                double lamL, double wcap, void* __restrict__ rbuf, int L, int K,
                                                int nts,
                                                const int* __restrict__ first_blk, int W) {
  __shared__ double2 rows[2][4][3][64];
  const int V3 = L * L * L;
  const int zspan = LINE ? L : NZ;  // z-sites per work-group
  const int nyb = W / 4, nzr = L / zspan;
  int ts, cb, g;
  unit_map(V3 / (4 * zspan), K / 64, nts, ts, cb, g);
  if (first_blk && cb < first_blk[ts]) return;  // whole work-group
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int zr = g % nzr;
  const int yb = (g / nzr) % nyb;
  const int t = g / nzr / nyb;
  const int x = t % L, strip = t / L;
  const int zbeg = zr * zspan;
  const int y = strip * W + yb * 4 + wave;
  const int l64 = threadIdx.x & 63;
  const double2* vb = v + (size_t)ts * V3 * 3 * K + cb * 64;  // column block of the timeslice
  const double2* Ut = U64 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int line = (x * L + y) * L;
  const int xup = x == L - 1 ? line - (L - 1) * L2 : line + L2;
  const int xdn = x == 0 ? line + (L - 1) * L2 : line - L2;
  const int yup = y == L - 1 ? line - (L - 1) * L : line + L;
  const int ydn = y == 0 ? line + (L - 1) * L : line - L;
  auto rw = [&](int site, int c) { return vb + (unsigned)((3 * site + c) * K); };
  double cs = 0.0, mc = 0.0;
  if (RES) {
    const double* m = mu + (size_t)ts * K;
    cs = window_scale(m, K, lamL, wcap);
    mc = m[cb * 64 + l64] - 6.0;
  }
  double2 zm[3], zc[3];
  {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      zm[c] = rw(line + (zbeg == 0 ? L - 1 : zbeg - 1), c)[l64];
      zc[c] = rw(line + zbeg, c)[l64];
    }
  }
  for (int z0 = zbeg; z0 < zbeg + zspan; z0 += NZ) {
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
      const int z = z0 + k;
      const int par = (z - zbeg) & 1;
      const int s = line + z;
      const int zdn = line + (z == 0 ? L - 1 : z - 1);
      const int zup = line + (z == L - 1 ? 0 : z + 1);
      double2 zn[3], yu[3], ym[3], xu[3], xm[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) rows[par][wave][c][l64] = zc[c];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zn[c] = rw(zup, c)[l64];
        xu[c] = rw(xup + z, c)[l64];
        xm[c] = rw(xdn + z, c)[l64];
      }
      if (wave == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yu[c] = rw(yup + z, c)[l64];
      }
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ym[c] = rw(ydn + z, c)[l64];
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (wave < 3) yu[c] = rows[par][wave + 1][c][l64];
        if (wave > 0) ym[c] = rows[par][wave - 1][c][l64];
      }
      const int su = __builtin_amdgcn_readfirstlane(s);
      const int dn[3] = {__builtin_amdgcn_readfirstlane(xdn + z), __builtin_amdgcn_readfirstlane(ydn + z),
                         __builtin_amdgcn_readfirstlane(zdn)};
      double2 acc[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) acc[c] = make_double2(0.0, 0.0);
#pragma unroll
      for (int mu_ = 0; mu_ < 3; ++mu_) {
        const double2* nu = mu_ == 0 ? xu : (mu_ == 1 ? yu : zn);
        const double2* nd = mu_ == 0 ? xm : (mu_ == 1 ? ym : zm);
        const double2* Uf = Ut + ((size_t)su * 3 + mu_) * 9;
        const double2* Ub = Ut + ((size_t)dn[mu_] * 3 + mu_) * 9;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
          for (int bb = 0; bb < 3; ++bb) {
            acc[a] = cfma(Uf[3 * a + bb], nu[bb], acc[a]);
            acc[a] = cfmac(Ub[3 * bb + a], nd[bb], acc[a]);
          }
        }
      }
      if (!RES) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          double2 o;
          o.x = acc[c].x - 6.0 * zc[c].x;
          o.y = acc[c].y - 6.0 * zc[c].y;
          st_nt(out + (size_t)ts * V3 * 3 * K + cb * 64 + (unsigned)((3 * su + c) * K) + l64, o);
        }
      } else {
        // block-major filter storage: block (cb*64 + l64)/BW, offset (cb*64 + l64) % BW within its rows
        // (a 64-column group lies in one block for BW >= 64, in two for BW = 32)
        const size_t n = (size_t)3 * V3;
        const int gc = cb * 64 + (BW < 64 ? l64 : 0);
        const size_t rbase = ((size_t)ts * (K / BW) + gc / BW) * n * BW + gc % BW;
        const int lo = BW < 64 ? 0 : l64;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float2 rv =
              make_float2((float)(cs * (acc[a].x + mc * zc[a].x)), (float)(cs * (acc[a].y + mc * zc[a].y)));
          const size_t ri = rbase + (unsigned)((3 * su + a) * BW) + lo;
          if (BF16)
            static_cast<unsigned*>(rbuf)[ri] = bf2_pack(rv);
          else
            static_cast<float2*>(rbuf)[ri] = rv;
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zm[c] = zc[c];
        zc[c] = zn[c];
      }
    }
  }
}

// part[ts][chunk][col] = |R|^2 summed over the kRedRows rows of one chunk,
// from r = c_t R in the filter storage, for the columns >= first_col.  grid
// (row chunks, active 64-column groups, ts); each thread sums every 4th row of
// one column, then the block adds its four row lanes in a fixed order.  The
// chunk partials are summed in chunk order by k_sum_chunks: no atomics, so a
// repeated solve reproduces every bit.
__global__ __launch_bounds__(256) void k_residual_norms(const void* __restrict__ rbuf,
                                                         const double* __restrict__ mu, double lamL, double wcap,
                                                         double* __restrict__ red, int n, int K,
                                                         int first_col, int bw, int bf16) {
  const int ts = blockIdx.z;
  const int col = first_col + blockIdx.y * 64 + (threadIdx.x & 63);
  const int r0 = blockIdx.x * kRedRows + (threadIdx.x >> 6);
  const int r1 = min(n, (int)(blockIdx.x + 1) * kRedRows);
  const double c = window_scale(mu + (size_t)ts * K, K, lamL, wcap);
  const size_t base = ((size_t)ts * (K / bw) + col / bw) * n * bw + col % bw;
  double acc = 0.0;
  for (int row = r0; row < r1; row += 4) {
    const size_t ri = base + (size_t)row * bw;
    const float2 rv = bf16 ? bf2_unpack(static_cast<const unsigned*>(rbuf)[ri]) : static_cast<const float2*>(rbuf)[ri];
    acc += (double)rv.x * rv.x + (double)rv.y * rv.y;
  }
  __shared__ double part[256];
  part[threadIdx.x] = acc / (c * c);
  __syncthreads();
  if (threadIdx.x < 64)
    red[((size_t)ts * gridDim.x + blockIdx.x) * K + col] =
        part[threadIdx.x] + part[threadIdx.x + 64] + part[threadIdx.x + 128] + part[threadIdx.x + 192];
}

// out[ts][col] = sum over chunk c = 0 .. nchunk-1 (in that order) of
// red[ts][c][col], for col >= first_col.
__global__ void k_sum_chunks(const double* __restrict__ red, double* __restrict__ out, int nts, int K, int first_col,
                             int nchunk) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = K - first_col;
  if (i >= nts * w) return;
  const int ts = i / w, col = first_col + i % w;
  const double* p = red + (size_t)ts * nchunk * K + col;
  double acc = 0.0;
  for (int c = 0; c < nchunk; ++c) acc += p[(size_t)c * K];
  out[(size_t)ts * K + col] = acc;
}

// Phases phi[ts][col] = arg X[ts][0][col] of the first nev columns
// (eigensystem.cpp:21), then X[ts][row][col] *= e^{-i phi} in place
// (eigensystem.cpp:22-27): two kernels, so row 0 is read before it changes.
__global__ void k_phases(const double2* __restrict__ X, double* __restrict__ phase, size_t per_ts, int K, int nev,
                         int nts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nts * nev) return;
  const int ts = i / nev, col = i % nev;
  const double2 x0 = X[(size_t)ts * per_ts + col];
  phase[i] = atan2(x0.y, x0.x);
}

__global__ void k_apply_phases(double2* __restrict__ X, const double* __restrict__ phase, size_t per_ts, int K,
                               int nev) {
  const int ts = blockIdx.y;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts) return;
  const int col = (int)(i % K);
  if (col >= nev) return;
  const double ph = phase[(size_t)ts * nev + col];
  const double cr = cos(ph), si = -sin(ph);
  const size_t e = (size_t)ts * per_ts + i;
  const double2 x = X[e];
  X[e] = make_double2(x.x * cr - x.y * si, x.x * si + x.y * cr);
}

// V[col][row] (column-major n x nev, the eigenvector-file layout) =
// X[row][col] for the first nev columns of one timeslice.
__global__ void k_extract_cols(const double2* __restrict__ X, double2* __restrict__ V, int n, int K) {
  const int col = blockIdx.y;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n; row += gridDim.x * blockDim.x)
    V[(size_t)col * n + row] = X[(size_t)row * K + col];
}

// U32c[ts][i] = c_t U32[ts][i], c_t = cab[ts].x: the filter's link copy for the
// current window of every timeslice of a lane (per_ts link entries each).
__global__ void k_scale_links32(const float2* __restrict__ U32, const float2* __restrict__ cab,
                                float2* __restrict__ U32c, size_t per_ts) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts) return;
  const size_t e = (size_t)blockIdx.y * per_ts + i;
  const float c = cab[blockIdx.y].x;
  const float2 u = U32[e];
  U32c[e] = make_float2(c * u.x, c * u.y);
}

hipError_t launch_scale_links32(const float2* U32, const float2* cab, float2* U32c, size_t per_ts, int nts,
                                hipStream_t st) {
  k_scale_links32<<<dim3((unsigned)((per_ts + 255) / 256), nts), 256, 0, st>>>(U32, cab, U32c, per_ts);
  return hipGetLastError();
}

__global__ void k_convert_links32(const double2* __restrict__ U, float2* __restrict__ U32, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) U32[i] = make_float2((float)U[i].x, (float)U[i].y);
}

// Seeded Gaussian start block (counter hash per element), FP64.
__global__ void k_random_block(double2* __restrict__ X, size_t total, unsigned long long seed) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  auto mix = [](unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const unsigned long long h1 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * i + 1) * 0xD1B54A32D192ED03ull);
  const unsigned long long h2 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * i + 2) * 0xD1B54A32D192ED03ull);
  // Box-Muller in FP32 hardware transcendentals: a start block only has to be
  // random and of full rank, and the FP64 library calls made this 25.8 GB
  // fill at 32^3 x 64 compute-bound (14 ms)
  const float u1 = ((float)(h1 >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((float)(h2 >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float rr = sqrtf(-2.0f * __logf(u1));
  float sn, cs;
  __sincosf(6.2831853f * u2, &sn, &cs);
  X[i] = make_double2((double)(rr * cs), (double)(rr * sn));
}

// Gaussian entries in columns [first_col, K) of the timeslices whose flag is
// non-zero (X[ts][row][K]); the others are left alone.
__global__ void k_reseed_cols(double2* __restrict__ X, int n, int K, int first_col, const int* __restrict__ flag,
                              unsigned long long seed) {
  const int ts = blockIdx.y;
  if (flag[ts] == 0) return;
  const int w = K - first_col;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * w) return;
  const size_t row = i / w, col = first_col + i % w;
  auto mix = [](unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const unsigned long long c = ((unsigned long long)ts * n + row) * K + col;
  const unsigned long long h1 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * c + 1) * 0xD1B54A32D192ED03ull);
  const unsigned long long h2 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * c + 2) * 0xD1B54A32D192ED03ull);
  const double u1 = ((double)(h1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)(h2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double rr = sqrt(-2.0 * log(u1));
  X[(size_t)ts * n * K + row * K + col] = make_double2(rr * cos(2.0 * M_PI * u2), rr * sin(2.0 * M_PI * u2));
}

// ---------------------------------------------------------------------------
static inline dim3 site_grid(int L, int K, int nts) {
  const int V3 = L * L * L;
  return dim3((unsigned)((V3 + 3) / 4) * (unsigned)(K / 64) * (unsigned)nts);
}

hipError_t launch_filter_step(const float2* U32, const void* in_, const void* r_, const void* old_, void* out_,
                              const float4* coef, const float2* cab, const int* first_blk, int L, int K, int nts,
                              int bw, int bf16, int W, int zrun, hipStream_t st) {
  if (bf16) {
    if (bw == 256)
      launch_bf16<4>(U32, in_, r_, old_, out_, coef, cab, first_blk, L, K, nts, W, st);
    else
      launch_bf16<2>(U32, in_, r_, old_, out_, coef, cab, first_blk, L, K, nts, W, st);
    return hipGetLastError();
  }
  const float2* in = static_cast<const float2*>(in_);
  const float2* r = static_cast<const float2*>(r_);
  const float2* old = static_cast<const float2*>(old_);
  float2* out = static_cast<float2*>(out_);
  if (bw == 128) {
    const int V3 = L * L * L;
    dim3 g((unsigned)((V3 + 3) / 4) * (unsigned)(K / 128) * (unsigned)nts);
    const float4* in4 = reinterpret_cast<const float4*>(in);
    const float4* r4 = reinterpret_cast<const float4*>(r);
    const float4* o4 = reinterpret_cast<const float4*>(old);
    float4* out4 = reinterpret_cast<float4*>(out);
    if (zrun > 0) {
      if (L % zrun != 0 || W % 4 != 0) return hipErrorInvalidValue;
      const dim3 gy((unsigned)(V3 / (4 * L)) * (unsigned)(K / 128) * (unsigned)nts);
#define LAPH_YGROUP_CASE(NZ)                                                                                   \
  case NZ:                                                                                                     \
    if (r && old)                                                                                              \
      k_filter_step32x2y<true, true, NZ><<<gy, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L, K, \
                                                             nts, W);                                          \
    else if (r)                                                                                                \
      k_filter_step32x2y<true, false, NZ><<<gy, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L,   \
                                                              K, nts, W);                                      \
    else if (old)                                                                                              \
      k_filter_step32x2y<false, true, NZ><<<gy, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L,   \
                                                              K, nts, W);                                      \
    else                                                                                                       \
      k_filter_step32x2y<false, false, NZ><<<gy, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L,  \
                                                               K, nts, W);                                     \
    return hipGetLastError();
      switch (zrun) {
        LAPH_YGROUP_CASE(1)
        LAPH_YGROUP_CASE(2)
        LAPH_YGROUP_CASE(4)
        LAPH_YGROUP_CASE(8)
        default:
          return hipErrorInvalidValue;
      }
#undef LAPH_YGROUP_CASE
    }
    if (r)
      k_filter_step32x2<true><<<g, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L, K, nts,
                                                 old != nullptr, W);
    else
      k_filter_step32x2<false><<<g, 256, 0, st>>>(U32, in4, r4, o4, out4, coef, cab, first_blk, L, K, nts,
                                                  old != nullptr, W);
    return hipGetLastError();
  }
  if (bw == 32) {
    const dim3 g8((unsigned)((L * L * L + 7) / 8) * (unsigned)(K / 32) * (unsigned)nts);
    if (r)
      k_filter_step32h<true><<<g8, 256, 0, st>>>(U32, in, r, old, out, coef, cab, first_blk, L, K, nts,
                                                  old != nullptr, W);
    else
      k_filter_step32h<false><<<g8, 256, 0, st>>>(U32, in, r, old, out, coef, cab, first_blk, L, K, nts,
                                                   old != nullptr, W);
    return hipGetLastError();
  }
  if (bw != 64) return hipErrorInvalidValue;
  if (zrun > 0) {
    const int V3 = L * L * L;
    if (L % zrun != 0 || W % 4 != 0) return hipErrorInvalidValue;
    const dim3 gy((unsigned)(V3 / (4 * L)) * (unsigned)(K / 64) * (unsigned)nts);
#define LAPH_Y64_CASE(NZ)                                                                                      \
  case NZ:                                                                                                     \
    if (r && old)                                                                                              \
      k_filter_step32y<true, true, NZ><<<gy, 256, 0, st>>>(U32, in, r, old, out, coef, first_blk, L, K, nts, W); \
    else if (r)                                                                                                \
      k_filter_step32y<true, false, NZ><<<gy, 256, 0, st>>>(U32, in, r, old, out, coef, first_blk, L, K, nts,    \
                                                            W);                                                \
    else if (old)                                                                                              \
      k_filter_step32y<false, true, NZ><<<gy, 256, 0, st>>>(U32, in, r, old, out, coef, first_blk, L, K, nts,    \
                                                            W);                                                \
    else                                                                                                       \
      k_filter_step32y<false, false, NZ><<<gy, 256, 0, st>>>(U32, in, r, old, out, coef, first_blk, L, K, nts,   \
                                                             W);                                               \
    return hipGetLastError();
    switch (zrun) {
      LAPH_Y64_CASE(1)
      LAPH_Y64_CASE(2)
      LAPH_Y64_CASE(4)
      LAPH_Y64_CASE(8)
      default:
        return hipErrorInvalidValue;
    }
#undef LAPH_Y64_CASE
  }
  if (r)
    k_filter_step32<true><<<site_grid(L, K, nts), 256, 0, st>>>(U32, in, r, old, out, coef, cab, first_blk, L, K, nts,
                                                              old != nullptr, W);
  else
    k_filter_step32<false><<<site_grid(L, K, nts), 256, 0, st>>>(U32, in, r, old, out, coef, cab, first_blk, L, K,
                                                               nts, old != nullptr, W);
  return hipGetLastError();
}

// D X walks one run of zrun sites per work-group (measured faster for it than
// whole lines), the residual kernel whole z-lines.
template <bool RES, int BW, bool BF16>
static hipError_t launch_lap64y(const double2* U64, const double2* v, double2* out, const double* mu, double lamL,
                                double wcap, void* r, int L, int K, int nts, const int* first_blk, int W, int zrun,
                                hipStream_t st) {
  const int zspan = RES ? L : zrun;
  const dim3 g((unsigned)(L * L * L / (4 * zspan)) * (unsigned)(K / 64) * (unsigned)nts);
  switch (zrun) {
    case 1:
      k_lap64y<RES, BW, BF16, 1, RES><<<g, 256, 0, st>>>(U64, v, out, mu, lamL, wcap, r, L, K, nts, first_blk, W);
      break;
    case 2:
      k_lap64y<RES, BW, BF16, 2, RES><<<g, 256, 0, st>>>(U64, v, out, mu, lamL, wcap, r, L, K, nts, first_blk, W);
      break;
    case 4:
      k_lap64y<RES, BW, BF16, 4, RES><<<g, 256, 0, st>>>(U64, v, out, mu, lamL, wcap, r, L, K, nts, first_blk, W);
      break;
    case 8:
      k_lap64y<RES, BW, BF16, 8, RES><<<g, 256, 0, st>>>(U64, v, out, mu, lamL, wcap, r, L, K, nts, first_blk, W);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_lap_block64(const double2* U64, const double2* v, double2* out, int L, int K, int nts,
                              const int* first_blk, int W, int zrun, hipStream_t st) {
  if (zrun > 0) {
    if (L % zrun != 0 || W % 4 != 0) return hipErrorInvalidValue;
    return launch_lap64y<false, 64, false>(U64, v, out, nullptr, 0.0, 0.0, nullptr, L, K, nts, first_blk, W, zrun,
                                                   st);
  }
  k_lap_block64<<<site_grid(L, K, nts), 256, 0, st>>>(U64, v, out, L, K, nts, first_blk, W);
  return hipGetLastError();
}

hipError_t launch_lap_residual64(const double2* U64, const double2* X, const double* mu, double lamL, double wcap,
                                 void* r, int L, int K, int nts, const int* first_blk, int W, int bw, int bf16,
                                 int zrun, hipStream_t st) {
  if (zrun > 0) {
    if (L % zrun != 0 || W % 4 != 0) return hipErrorInvalidValue;
    if (bf16 && bw == 256)
      return launch_lap64y<true, 256, true>(U64, X, nullptr, mu, lamL, wcap, r, L, K, nts, first_blk, W, zrun, st);
    if (bf16 && bw == 128)
      return launch_lap64y<true, 128, true>(U64, X, nullptr, mu, lamL, wcap, r, L, K, nts, first_blk, W, zrun, st);
    if (!bf16 && bw == 128)
      return launch_lap64y<true, 128, false>(U64, X, nullptr, mu, lamL, wcap, r, L, K, nts, first_blk, W, zrun, st);
    if (!bf16 && bw == 64)
      return launch_lap64y<true, 64, false>(U64, X, nullptr, mu, lamL, wcap, r, L, K, nts, first_blk, W, zrun, st);
    if (!bf16 && bw == 32)
      return launch_lap64y<true, 32, false>(U64, X, nullptr, mu, lamL, wcap, r, L, K, nts, first_blk, W, zrun, st);
    return hipErrorInvalidValue;
  }
  k_lap_residual64<<<site_grid(L, K, nts), 256, 0, st>>>(U64, X, mu, lamL, wcap, r, L, K, nts, first_blk, W, bw,
                                                         bf16);
  return hipGetLastError();
}

hipError_t launch_residual_norms(const void* r, const double* mu, double lamL, double wcap, double* nrm2,
                                 double* red, int n, int K, int nts, int first_col, int bw, int bf16,
                                 hipStream_t st) {
  const int w = K - first_col;
  const int nchunk = red_chunks(n);
  dim3 g(nchunk, w / 64, nts);
  k_residual_norms<<<g, 256, 0, st>>>(r, mu, lamL, wcap, red, n, K, first_col, bw, bf16);
  k_sum_chunks<<<(nts * w + 255) / 256, 256, 0, st>>>(red, nrm2, nts, K, first_col, nchunk);
  return hipGetLastError();
}

hipError_t launch_fix_phases(double2* X, double* phase, int n, int K, int nev, int nts, hipStream_t st) {
  const size_t per_ts = (size_t)n * K;
  k_phases<<<(nts * nev + 255) / 256, 256, 0, st>>>(X, phase, per_ts, K, nev, nts);
  k_apply_phases<<<dim3((unsigned)((per_ts + 255) / 256), nts), 256, 0, st>>>(X, phase, per_ts, K, nev);
  return hipGetLastError();
}

hipError_t launch_extract_cols(const double2* X, double2* V, int n, int K, int nev, hipStream_t st) {
  k_extract_cols<<<dim3(64, nev), 256, 0, st>>>(X, V, n, K);
  return hipGetLastError();
}

hipError_t launch_convert_links32(const double2* U, float2* U32, size_t n, hipStream_t st) {
  k_convert_links32<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(U, U32, n);
  return hipGetLastError();
}

hipError_t launch_reseed_cols(double2* X, int n, int K, int nts, int first_col, const int* flag,
                              unsigned long long seed, hipStream_t st) {
  const size_t total = (size_t)n * (K - first_col);
  k_reseed_cols<<<dim3((unsigned)((total + 255) / 256), nts), 256, 0, st>>>(X, n, K, first_col, flag, seed);
  return hipGetLastError();
}

hipError_t launch_random_block(double2* X, size_t total, unsigned long long seed, hipStream_t st) {
  k_random_block<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(X, total, seed);
  return hipGetLastError();
}

}  // namespace laph

namespace laph {

// red[ts][chunk][col] = sum over the kRedRows rows of one chunk of
// |Y[ts][row][col]|^2 for col >= first_col (summed in chunk order by
// k_sum_chunks, as the residual norms).
__global__ __launch_bounds__(256) void k_colnorm2_blk(const double2* __restrict__ Y, double* __restrict__ red, int n,
                                                       int K, int first_col) {
  const int ts = blockIdx.z;
  const int col = first_col + blockIdx.y * 64 + (threadIdx.x & 63);
  const int r0 = blockIdx.x * kRedRows + (threadIdx.x >> 6), r1 = min(n, (int)(blockIdx.x + 1) * kRedRows);
  const size_t off = (size_t)ts * n * K;
  double acc = 0.0;
  for (int r = r0; r < r1; r += 4) {
    const double2 v = Y[off + (size_t)r * K + col];
    acc += v.x * v.x + v.y * v.y;
  }
  __shared__ double part[256];
  part[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64)
    red[((size_t)ts * gridDim.x + blockIdx.x) * K + col] =
        part[threadIdx.x] + part[threadIdx.x + 64] + part[threadIdx.x + 128] + part[threadIdx.x + 192];
}

// Y[ts][row][col] /= sqrt(nrm[ts][col]) for col >= first_col.
__global__ void k_scale_cols(double2* __restrict__ Y, const double* __restrict__ nrm, size_t per_ts, int K,
                             int first_col) {
  const int ts = blockIdx.y;
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts) return;
  const int col = (int)(i % K);
  if (col < first_col) return;
  const double nv = nrm[(size_t)ts * K + col];
  const double s = nv > 0.0 ? 1.0 / sqrt(nv) : 1.0;
  double2 v = Y[(size_t)ts * per_ts + i];
  Y[(size_t)ts * per_ts + i] = make_double2(v.x * s, v.y * s);
}

hipError_t launch_normalize_columns(double2* Y, double* tmp, double* red, int n, int K, int nts, int first_col,
                                    hipStream_t st) {
  const int w = K - first_col;
  const int nchunk = red_chunks(n);
  dim3 g(nchunk, w / 64, nts);
  k_colnorm2_blk<<<g, 256, 0, st>>>(Y, red, n, K, first_col);
  k_sum_chunks<<<(nts * w + 255) / 256, 256, 0, st>>>(red, tmp, nts, K, first_col, nchunk);
  const size_t per_ts = (size_t)n * K;
  dim3 gs((unsigned)((per_ts + 255) / 256), nts);
  k_scale_cols<<<gs, 256, 0, st>>>(Y, tmp, per_ts, K, first_col);
  return hipGetLastError();
}

}  // namespace laph
