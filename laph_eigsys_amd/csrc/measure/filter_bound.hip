// What bounds the FP32 filter step: the library's step (k_filter_step32x2y:
// 128-vector blocks, work-groups of 4 y-rows walking z-lines in runs of 8, one
// correction-pass factor, every column active) timed beside variants of the
// same kernel with one part taken away, and beside copies that keep its grid,
// its site order and its load pattern but do no link products:
//
//   step                      the library's kernel (launch_filter_step)
//   step_mirror               this file's restatement of it (bitwise equal)
//   step_no_link_products     the six 3x3 link products replaced by adds
//   step_neighbours_cached    every neighbour row read from the wavefront's
//                             own line instead (L1/L2 hits, same count)
//   copy_zwalk                z-walk: read the own line once, write it
//   copy_zwalk_lds            + the LDS row exchange and per-site barrier
//   copy_zwalk_lds_x          + the two x-neighbour rows of every site
//   copy_zwalk_lds_x_y        + the y-edge rows of waves 0 and 3 (the step's
//                             full load pattern)
//   copy_zwalk_lds_x_ycached  the y-edge rows taken from the own line
//   stream_copy               grid-stride float4 copy (1024 work-groups)
//
// Prints one JSON object: microseconds per launch (mean of 10 after a warm-up
// round) and GB/s of the step's algorithmic bytes (2 x 24 B per site and
// vector + 216 B of links per site and block) for each; the copies move
// 48 B per site and vector.  Usage: filter_bound [L [timeslices [strip]]].
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#include "../kernels_block.h"

#define CK(x)                                                                            \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void st_nt4(float4* p, float4 v) {
  f4v w;
  w.x = v.x;
  w.y = v.y;
  w.z = v.z;
  w.w = v.w;
  __builtin_nontemporal_store(w, reinterpret_cast<f4v*>(p));
}
__device__ __forceinline__ float2 cfma(float2 u, float2 v, float2 acc) {
  acc.x = fma(u.x, v.x, fma(-u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(u.y, v.x, acc.y));
  return acc;
}
__device__ __forceinline__ float2 cfmac(float2 u, float2 v, float2 acc) {
  acc.x = fma(u.x, v.x, fma(u.y, v.y, acc.x));
  acc.y = fma(u.x, v.y, fma(-u.y, v.x, acc.y));
  return acc;
}
// the library's grid decode: every XCD (linear id mod 8) owns whole units
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
__device__ __forceinline__ const float4* row(const float4* p, int site, int c) {
  return p + (unsigned)((3 * site + c) * 64);
}
template <bool LINKS>
__device__ __forceinline__ void terms(const float2* __restrict__ Uf, const float2* __restrict__ Ub, const float4 xu[3],
                                      const float4 xm[3], float2 acc0[3], float2 acc1[3]) {
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    if (!LINKS) {
      acc0[a].x += xu[a].x + xm[a].x;
      acc0[a].y += xu[a].y + xm[a].y;
      acc1[a].x += xu[a].z + xm[a].z;
      acc1[a].y += xu[a].w + xm[a].w;
      continue;
    }
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

// The step with parts switchable.  LINKS: the link products (else adds).
// NBR: 0 the neighbour rows, 1 the same number of rows from the own line.
template <bool LINKS, int NBR>
__global__ __launch_bounds__(256) void k_step(const float2* __restrict__ U32, const float4* __restrict__ in,
                                              float4* __restrict__ out, const float4* __restrict__ coef, int L,
                                              int K, int nts, int W) {
  __shared__ float4 rows[2][4][3][64];
  const int V3 = L * L * L;
  const int nyb = W / 4;
  int ts, cb, g;
  unit_map(V3 / (4 * L), K / 128, nts, ts, cb, g);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int yb = g % nyb;
  const int t = g / nyb;
  const int x = t % L, strip = t / L;
  const int y = strip * W + yb * 4 + wave;
  const int l64 = threadIdx.x & 63;
  const int col = cb * 128 + 2 * l64;
  const size_t vo = ((size_t)ts * (K / 128) + cb) * (size_t)V3 * 3 * 64;
  const float2* Ut = U32 + (size_t)ts * V3 * 27;
  const int L2 = L * L;
  const int line = (x * L + y) * L;
  const int xup = x == L - 1 ? line - (L - 1) * L2 : line + L2;
  const int xdn = x == 0 ? line + (L - 1) * L2 : line - L2;
  const int yup = y == L - 1 ? line - (L - 1) * L : line + L;
  const int ydn = y == 0 ? line + (L - 1) * L : line - L;
  const float4* v = in + vo;
  const float4 cf0 = coef[(size_t)ts * K + col], cf1 = coef[(size_t)ts * K + col + 1];
  float4 zm[3], zc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zm[c] = row(v, line + L - 1, c)[l64];
    zc[c] = row(v, line, c)[l64];
  }
  for (int z0 = 0; z0 < L; z0 += 8) {
#pragma unroll
    for (int k = 0; k < 8; ++k) {
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
        xu[c] = row(v, NBR ? s : xup + z, c)[l64];
        xm[c] = row(v, NBR ? zdn : xdn + z, c)[l64];
      }
      if (wave == 3) {
#pragma unroll
        for (int c = 0; c < 3; ++c) yu[c] = row(v, NBR ? s : yup + z, c)[l64];
      }
      if (wave == 0) {
#pragma unroll
        for (int c = 0; c < 3; ++c) ym[c] = row(v, NBR ? zdn : ydn + z, c)[l64];
      }
      float2 acc0[3], acc1[3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        acc0[c] = make_float2(0.f, 0.f);
        acc1[c] = make_float2(0.f, 0.f);
      }
      const int su = __builtin_amdgcn_readfirstlane(s), xd = __builtin_amdgcn_readfirstlane(xdn + z);
      const int yd = __builtin_amdgcn_readfirstlane(ydn + z), zd = __builtin_amdgcn_readfirstlane(zdn);
      terms<LINKS>(Ut + ((size_t)su * 3 + 0) * 9, Ut + ((size_t)xd * 3 + 0) * 9, xu, xm, acc0, acc1);
      terms<LINKS>(Ut + ((size_t)su * 3 + 2) * 9, Ut + ((size_t)zd * 3 + 2) * 9, zn, zm, acc0, acc1);
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        if (wave < 3) yu[c] = rows[par][wave + 1][c][l64];
        if (wave > 0) ym[c] = rows[par][wave - 1][c][l64];
      }
      terms<LINKS>(Ut + ((size_t)su * 3 + 1) * 9, Ut + ((size_t)yd * 3 + 1) * 9, yu, ym, acc0, acc1);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float4 ctr = zc[c];
        float4 o;
        o.x = cf0.x * acc0[c].x + cf0.w * ctr.x;
        o.y = cf0.x * acc0[c].y + cf0.w * ctr.y;
        o.z = cf1.x * acc1[c].x + cf1.w * ctr.z;
        o.w = cf1.x * acc1[c].y + cf1.w * ctr.w;
        o.x -= cf0.z * ctr.y;
        o.y += cf0.z * ctr.x;
        o.z -= cf1.z * ctr.w;
        o.w += cf1.z * ctr.z;
        st_nt4(out + vo + (unsigned)((3 * su + c) * 64) + l64, o);
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        zm[c] = zc[c];
        zc[c] = zn[c];
      }
    }
  }
}

// Copies over the step's grid and z-walk.  CM: 0 plain; 1 + LDS exchange and
// barrier; 2 + x-neighbour rows; 3 + y-edge rows (waves 0 and 3); 4 y-edge
// rows from the own line.
template <int CM>
__global__ __launch_bounds__(256) void k_zcopy(const float4* __restrict__ in, float4* __restrict__ out, int L, int K,
                                               int nts, int W) {
  __shared__ float4 rows[2][4][3][64];
  const int V3 = L * L * L;
  const int nyb = W / 4;
  int ts, cb, g;
  unit_map(V3 / (4 * L), K / 128, nts, ts, cb, g);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int yb = g % nyb;
  const int t = g / nyb;
  const int x = t % L, strip = t / L;
  const int y = strip * W + yb * 4 + wave;
  const int l64 = threadIdx.x & 63;
  const size_t vo = ((size_t)ts * (K / 128) + cb) * (size_t)V3 * 3 * 64;
  const int L2 = L * L;
  const int line = (x * L + y) * L;
  const int xup = x == L - 1 ? line - (L - 1) * L2 : line + L2;
  const int xdn = x == 0 ? line + (L - 1) * L2 : line - L2;
  const int yedge = CM == 4 ? line : wave == 0 ? (y == 0 ? line + (L - 1) * L : line - L)
                                              : (y == L - 1 ? line - (L - 1) * L : line + L);
  const float4* v = in + vo;
  float4 zc[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) zc[c] = row(v, line, c)[l64];
  for (int z = 0; z < L; ++z) {
    const int par = z & 1;
    const int s = line + z;
    float4 nx[3], o[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      nx[c] = row(v, line + (z == L - 1 ? 0 : z + 1), c)[l64];
      o[c] = zc[c];
    }
    if (CM >= 1) {
#pragma unroll
      for (int c = 0; c < 3; ++c) rows[par][wave][c][l64] = zc[c];
      if (CM >= 2) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float4 a = row(v, xup + z, c)[l64], b = row(v, xdn + z, c)[l64];
          o[c].x += a.x + b.x;
          o[c].y += a.y + b.y;
          o[c].z += a.z + b.z;
          o[c].w += a.w + b.w;
        }
      }
      if (CM >= 3 && (wave == 0 || wave == 3)) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float4 a = row(v, yedge + z, c)[l64];
          o[c].x += a.x;
          o[c].y += a.y;
          o[c].z += a.z;
          o[c].w += a.w;
        }
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float4 a = rows[par][(wave + 1) & 3][c][l64];
        o[c].x += a.x;
        o[c].y += a.y;
        o[c].z += a.z;
        o[c].w += a.w;
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      st_nt4(out + vo + (unsigned)((3 * s + c) * 64) + l64, o[c]);
      zc[c] = nx[c];
    }
  }
}

__global__ __launch_bounds__(256) void k_stream(const float4* __restrict__ a, float4* __restrict__ o, size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const f4v v = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(a + i));
    __builtin_nontemporal_store(v, reinterpret_cast<f4v*>(o + i));
  }
}

__global__ void k_fill(float* p, size_t n, unsigned seed) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    unsigned h = (unsigned)i * 2654435761u ^ seed;
    h ^= h >> 15;
    h *= 2246822519u;
    h ^= h >> 13;
    h *= 3266489917u;
    h ^= h >> 16;
    p[i] = (float)(h & 0xffffff) * (1.0f / 16777216.0f) - 0.5f;
  }
}

__global__ void k_count_diff(const unsigned* a, const unsigned* b, size_t n, unsigned long long* cnt) {
  unsigned long long c = 0;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) c += a[i] != b[i];
  if (c) atomicAdd(cnt, c);
}

}  // namespace

int main(int argc, char** argv) {
  const int L = argc > 1 ? std::atoi(argv[1]) : 32;
  const int nts = argc > 2 ? std::atoi(argv[2]) : 64;
  const int W = argc > 3 ? std::atoi(argv[3]) : 4;
  const int K = 256, reps = 10;
  if (L < 8 || L % 8 != 0 || nts < 1 || W % 4 != 0 || L % W != 0) {
    std::fprintf(stderr, "filter_bound: need L a multiple of 8, timeslices >= 1, strip a multiple of 4 dividing L\n");
    return 2;
  }
  const size_t V3 = (size_t)L * L * L;
  const size_t nvec = (size_t)nts * V3 * 3 * K * 2;  // floats per vector buffer
  const size_t nlink = (size_t)nts * V3 * 27 * 2;
  float *in, *out, *out2, *U, *coef;
  int* fb;
  float2* cab;
  unsigned long long* cnt;
  CK(hipMalloc(&in, nvec * 4));
  CK(hipMalloc(&out, nvec * 4));
  CK(hipMalloc(&out2, nvec * 4));
  CK(hipMalloc(&U, nlink * 4));
  CK(hipMalloc(&coef, (size_t)nts * K * 16));
  CK(hipMalloc(&fb, nts * sizeof(int)));
  CK(hipMalloc(&cab, nts * sizeof(float2)));
  CK(hipMalloc(&cnt, sizeof(unsigned long long)));
  CK(hipMemset(fb, 0, nts * sizeof(int)));
  CK(hipMemset(cab, 0, nts * sizeof(float2)));
  CK(hipMemset(cnt, 0, sizeof(unsigned long long)));
  k_fill<<<4096, 256>>>(in, nvec, 1);
  k_fill<<<4096, 256>>>(U, nlink, 2);
  k_fill<<<64, 256>>>(coef, (size_t)nts * K * 4, 3);
  CK(hipGetLastError());
  CK(hipDeviceSynchronize());
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const double alg = (double)nts * V3 * K * (48.0 + 216.0 / 128.0);
  const dim3 grid((unsigned)(V3 / (4 * L)) * (K / 128) * nts);
  const float2* U2 = reinterpret_cast<const float2*>(U);
  const float4* in4 = reinterpret_cast<const float4*>(in);
  const float4* cf = reinterpret_cast<const float4*>(coef);
  float4* o2 = reinterpret_cast<float4*>(out2);
  std::string js = "{";
  char buf[256];
  std::snprintf(buf, sizeof buf, "\"L\": %d, \"timeslices\": %d, \"vectors\": %d, \"strip\": %d, "
                "\"alg_bytes_per_launch\": %.0f", L, nts, K, W, alg);
  js += buf;
  auto timeit = [&](const char* name, auto launch) {
    float ms = 0;
    for (int round = 0; round < 2; ++round) {  // the first round warms up
      CK(hipEventRecord(e0));
      for (int r = 0; r < reps; ++r) launch();
      CK(hipGetLastError());
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
    }
    ms /= reps;
    std::snprintf(buf, sizeof buf, ", \"%s_us\": %.1f, \"%s_GBps\": %.0f", name, ms * 1e3, name,
                  alg / (ms * 1e-3) / 1e9);
    js += buf;
  };
  timeit("step", [&] {
    CK(laph::launch_filter_step(U2, in, nullptr, nullptr, out, cf, cab, fb, L, K, nts, 128, 0, W, 8, 0));
  });
  timeit("step_mirror", [&] { k_step<true, 0><<<grid, 256>>>(U2, in4, o2, cf, L, K, nts, W); });
  k_count_diff<<<4096, 256>>>(reinterpret_cast<const unsigned*>(out), reinterpret_cast<const unsigned*>(out2), nvec,
                              cnt);
  CK(hipGetLastError());
  unsigned long long ndiff = 0;
  CK(hipMemcpy(&ndiff, cnt, sizeof ndiff, hipMemcpyDeviceToHost));
  std::snprintf(buf, sizeof buf, ", \"step_mirror_words_differing\": %llu", ndiff);
  js += buf;
  timeit("step_no_link_products", [&] { k_step<false, 0><<<grid, 256>>>(U2, in4, o2, cf, L, K, nts, W); });
  timeit("step_neighbours_cached", [&] { k_step<true, 1><<<grid, 256>>>(U2, in4, o2, cf, L, K, nts, W); });
  timeit("copy_zwalk", [&] { k_zcopy<0><<<grid, 256>>>(in4, o2, L, K, nts, W); });
  timeit("copy_zwalk_lds", [&] { k_zcopy<1><<<grid, 256>>>(in4, o2, L, K, nts, W); });
  timeit("copy_zwalk_lds_x", [&] { k_zcopy<2><<<grid, 256>>>(in4, o2, L, K, nts, W); });
  timeit("copy_zwalk_lds_x_y", [&] { k_zcopy<3><<<grid, 256>>>(in4, o2, L, K, nts, W); });
  timeit("copy_zwalk_lds_x_ycached", [&] { k_zcopy<4><<<grid, 256>>>(in4, o2, L, K, nts, W); });
  timeit("stream_copy", [&] { k_stream<<<1024, 256>>>(in4, o2, nvec / 4); });
  js += "}";
  std::printf("%s\n", js.c_str());
  return 0;
}
