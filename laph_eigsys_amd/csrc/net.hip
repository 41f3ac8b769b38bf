// Learned warm start: a gauge-covariant parallel-transport network that maps
// Gaussian random colour vectors to a starting block for the batched solver.
//
// Layer l (C = 16 complex channels per group, K/16 groups per timeslice):
//   z_c(x)  = sum_t sum_c' W_l[t][c][c'] (T_t a_c')(x)
//   T_0 psi(x)        = psi(x)
//   T_{1+2mu} psi(x)  = U_mu(x) psi(x+mu)
//   T_{2+2mu} psi(x)  = U_mu(x-mu)^H psi(x-mu)          (mu = x, y, z)
//   a_c(x)  = z_c(x) * max(0, |z_c(x)| + b_c) / |z_c(x)|  (all but the last layer)
// |z_c(x)| is the colour norm, so every layer is gauge covariant:
// psi(x) -> Omega(x) psi(x), U_mu(x) -> Omega(x) U_mu(x) Omega(x+mu)^H.
// Activations are stored in BF16 (or FP32 for the gradient check), all
// arithmetic is FP32.
//
// Training maximises the fraction of each output vector inside the span of
// the solver's converged eigenvectors V:
//   loss = -(1/Kc) sum_j |V^H y_j|^2 / |y_j|^2,
// whose conjugate gradient is g_j = -(1/Kc) (V V^H y_j - f_j y_j) / |y_j|^2.
// Backward through a layer uses the same fused kernel with conj(W)^T and the
// adjoint hops (T_{1+2mu}^H = T_{2+2mu}); weight gradients are per-block
// partial sums reduced by a second kernel.  Gradients returned to the caller
// are real gradients of (Re W, Im W, b).
#include <rocblas/rocblas.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "batch_internal.h"
#include "kernels_block.h"

namespace laph {

constexpr int NC = 16;   // channels per group
constexpr int NT = 7;    // self + 6 hops
constexpr int WSZ = NT * NC * NC;  // complex weights per layer

struct bf16c {
  unsigned short x, y;
};

__device__ __forceinline__ unsigned short f2bf(float f) {  // round to nearest even
  unsigned u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(((unsigned)h) << 16); }

template <class A>
struct act_io;
template <>
struct act_io<float2> {
  static __device__ __forceinline__ float2 ld(const float2* p, size_t i) { return p[i]; }
  static __device__ __forceinline__ void st(float2* p, size_t i, float2 v) { p[i] = v; }
};
template <>
struct act_io<bf16c> {
  static __device__ __forceinline__ float2 ld(const bf16c* p, size_t i) {
    const bf16c v = p[i];
    return make_float2(bf2f(v.x), bf2f(v.y));
  }
  static __device__ __forceinline__ void st(bf16c* p, size_t i, float2 v) { p[i] = bf16c{f2bf(v.x), f2bf(v.y)}; }
};

__device__ __forceinline__ float2 cmaf(float2 u, float2 v, float2 acc) {  // acc += u*v
  acc.x = fmaf(u.x, v.x, fmaf(-u.y, v.y, acc.x));
  acc.y = fmaf(u.x, v.y, fmaf(u.y, v.x, acc.y));
  return acc;
}
__device__ __forceinline__ float2 cmafc(float2 u, float2 v, float2 acc) {  // acc += conj(u)*v
  acc.x = fmaf(u.x, v.x, fmaf(u.y, v.y, acc.x));
  acc.y = fmaf(u.x, v.y, fmaf(-u.y, v.x, acc.y));
  return acc;
}

__device__ __forceinline__ void neighbours(int s, int L, int up[3], int dn[3]) {
  const int L2 = L * L;
  const int x = s / L2, y = (s / L) % L, z = s % L;
  up[0] = x == L - 1 ? s - (L - 1) * L2 : s + L2;
  up[1] = y == L - 1 ? s - (L - 1) * L : s + L;
  up[2] = z == L - 1 ? s - (L - 1) : s + 1;
  dn[0] = x == 0 ? s + (L - 1) * L2 : s - L2;
  dn[1] = y == 0 ? s + (L - 1) * L : s - L;
  dn[2] = z == 0 ? s + (L - 1) : s - 1;
}

// The 7 transported copies of the lane's column at site s into T[t][colour].
template <class A>
__device__ __forceinline__ void transports(const A* __restrict__ in, const float2* __restrict__ U, int s, int L, int K,
                                           int col, float2 T[NT][3]) {
  int up[3], dn[3];
  neighbours(s, L, up, dn);
#pragma unroll
  for (int c = 0; c < 3; ++c) T[0][c] = act_io<A>::ld(in, (size_t)(3 * s + c) * K + col);
#pragma unroll
  for (int mu = 0; mu < 3; ++mu) {
    float2 pu[3], pd[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      pu[c] = act_io<A>::ld(in, (size_t)(3 * up[mu] + c) * K + col);
      pd[c] = act_io<A>::ld(in, (size_t)(3 * dn[mu] + c) * K + col);
    }
    const float2* Uf = U + ((size_t)s * 3 + mu) * 9;       // wave-uniform
    const float2* Ub = U + ((size_t)dn[mu] * 3 + mu) * 9;  // wave-uniform
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float2 f = make_float2(0.f, 0.f), g = make_float2(0.f, 0.f);
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        f = cmaf(Uf[3 * a + b], pu[b], f);
        g = cmafc(Ub[3 * b + a], pd[b], g);
      }
      T[1 + 2 * mu][a] = f;
      T[2 + 2 * mu][a] = g;
    }
  }
}

__device__ __forceinline__ int swap_hop(int t) { return t == 0 ? 0 : ((t & 1) ? t + 1 : t - 1); }

// Fused transport + channel mixing (+ covariant gate).  One wavefront = one
// site x 64 columns (4 groups of 16 channels); 4 wavefronts per block.
// ADJ: apply the adjoint layer (conj(W)^T, adjoint hops) for backprop.
template <class AIN, class AOUT, bool ADJ>
__global__ __launch_bounds__(256) void k_net_mix(const float2* __restrict__ W, const float* __restrict__ bias,
                                                  const AIN* __restrict__ in, AOUT* __restrict__ out,
                                                  AOUT* __restrict__ zout, const float2* __restrict__ U32, int L,
                                                  int K, size_t per_ts) {
  __shared__ float2 Wsh[NT][NC][NC];
  __shared__ float2 Tsh[4][NT][3][64];
  const int tid = threadIdx.x;
  for (int i = tid; i < WSZ; i += 256) {
    const int t = i / (NC * NC), c = (i / NC) % NC, cp = i % NC;
    if (ADJ) {
      const float2 w = W[(swap_hop(t) * NC + cp) * NC + c];
      Wsh[t][c][cp] = make_float2(w.x, -w.y);
    } else {
      Wsh[t][c][cp] = W[i];
    }
  }
  const int ts = blockIdx.z;
  const int V3 = L * L * L;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int s = blockIdx.x * 4 + wave;
  const bool live = s < V3;
  const int col = blockIdx.y * 64 + lane;
  const AIN* inT = in + per_ts * ts;
  float2 T[NT][3];
  if (live) transports<AIN>(inT, U32 + (size_t)ts * V3 * 27, s, L, K, col, T);
#pragma unroll
  for (int t = 0; t < NT; ++t)
#pragma unroll
    for (int c = 0; c < 3; ++c) Tsh[wave][t][c][lane] = live ? T[t][c] : make_float2(0.f, 0.f);
  __syncthreads();
  if (!live) return;
  const int g0 = lane & ~(NC - 1), c = lane & (NC - 1);
  float2 o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
  for (int t = 0; t < NT; ++t)
#pragma unroll 4
    for (int cp = 0; cp < NC; ++cp) {
      const float2 w = Wsh[t][c][cp];
#pragma unroll
      for (int k = 0; k < 3; ++k) o[k] = cmaf(w, Tsh[wave][t][k][g0 + cp], o[k]);
    }
  AOUT* outT = out + per_ts * ts;
  if (bias) {
    AOUT* zT = zout ? zout + per_ts * ts : nullptr;
    const float r = sqrtf(o[0].x * o[0].x + o[0].y * o[0].y + o[1].x * o[1].x + o[1].y * o[1].y +
                          o[2].x * o[2].x + o[2].y * o[2].y);
    const float b = bias[c];
    const float sc = (r > 0.f && r + b > 0.f) ? (r + b) / r : 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (zout) act_io<AOUT>::st(zT, (size_t)(3 * s + k) * K + col, o[k]);
      act_io<AOUT>::st(outT, (size_t)(3 * s + k) * K + col, make_float2(o[k].x * sc, o[k].y * sc));
    }
  } else {
#pragma unroll
    for (int k = 0; k < 3; ++k) act_io<AOUT>::st(outT, (size_t)(3 * s + k) * K + col, o[k]);
  }
}

// Gate backward: g_z = (1 + b/r) g_y - (b/r^3) Re<g_y, z> z where active;
// bias partials sum 2 Re<g_y, z>/r per (block, channel).
template <class A>
__global__ __launch_bounds__(256) void k_net_gate_bwd(const A* __restrict__ z, const float* __restrict__ bias,
                                                       const float2* __restrict__ gy, float2* __restrict__ gz,
                                                       float* __restrict__ bpart, int V3, int K, size_t per_ts) {
  __shared__ float red[256];
  const int ts = blockIdx.z;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int col = blockIdx.y * 64 + (threadIdx.x & 63);
  const int c = col & (NC - 1);
  float contrib = 0.f;
  if (s < V3) {
    const size_t o = per_ts * ts + (size_t)3 * s * K + col;
    float2 zv[3], gv[3];
    float r2 = 0.f, dot = 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      zv[k] = act_io<A>::ld(z, o + (size_t)k * K);
      gv[k] = gy[o + (size_t)k * K];
      r2 += zv[k].x * zv[k].x + zv[k].y * zv[k].y;
      dot += gv[k].x * zv[k].x + gv[k].y * zv[k].y;
    }
    const float r = sqrtf(r2), b = bias[c];
    const bool on = r > 0.f && r + b > 0.f;
    const float s1 = on ? 1.f + b / r : 0.f;
    const float s2 = on ? -b / (r * r2) * dot : 0.f;
#pragma unroll
    for (int k = 0; k < 3; ++k)
      gz[o + (size_t)k * K] = make_float2(s1 * gv[k].x + s2 * zv[k].x, s1 * gv[k].y + s2 * zv[k].y);
    contrib = on ? 2.f * dot / r : 0.f;
  }
  red[threadIdx.x] = contrib;
  __syncthreads();
  if (threadIdx.x < NC) {  // channel c = lane & 15: sum the 16 lanes of every group and the 4 waves
    float acc = 0.f;
    for (int i = threadIdx.x; i < 256; i += NC) acc += red[i];
    const size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
    bpart[blk * NC + threadIdx.x] = acc;
  }
}

// Weight-gradient partials: for every block, sum over its sites of
//   gw[t][c][c'] += sum_{groups, colour} g_c(x) conj((T_t a_c')(x)).
// Each block sweeps SITES_PER_BLOCK sites, 4 at a time (one per wavefront).
constexpr int SITES_PER_BLOCK = 64;
template <class A>
__global__ __launch_bounds__(256) void k_net_wgrad(const A* __restrict__ a, const float2* __restrict__ g,
                                                    const float2* __restrict__ U32, float2* __restrict__ part, int L,
                                                    int K, size_t per_ts) {
  __shared__ float2 Tsh[4][NT][3][64];
  __shared__ float2 Gsh[4][3][64];
  const int ts = blockIdx.z;
  const int V3 = L * L * L;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int col = blockIdx.y * 64 + lane;
  const int c = threadIdx.x >> 4, cp = threadIdx.x & (NC - 1);
  float2 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = make_float2(0.f, 0.f);
  const A* aT = a + per_ts * ts;
  const float2* gT = g + per_ts * ts;
  const float2* UT = U32 + (size_t)ts * V3 * 27;
  for (int base = blockIdx.x * SITES_PER_BLOCK; base < min(V3, (int)(blockIdx.x + 1) * SITES_PER_BLOCK); base += 4) {
    const int s = base + wave;
    float2 T[NT][3];
    if (s < V3) {
      transports<A>(aT, UT, s, L, K, col, T);
#pragma unroll
      for (int k = 0; k < 3; ++k) Gsh[wave][k][lane] = gT[(size_t)(3 * s + k) * K + col];
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int k = 0; k < 3; ++k) T[t][k] = make_float2(0.f, 0.f);
#pragma unroll
      for (int k = 0; k < 3; ++k) Gsh[wave][k][lane] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int k = 0; k < 3; ++k) Tsh[wave][t][k][lane] = T[t][k];
    __syncthreads();
    for (int w = 0; w < 4; ++w)
#pragma unroll
      for (int gg = 0; gg < 64; gg += NC)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float2 gv = Gsh[w][k][gg + c];
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[t] = cmafc(Tsh[w][t][k][gg + cp], gv, acc[t]);
        }
    __syncthreads();
  }
  const size_t blk = ((size_t)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
#pragma unroll
  for (int t = 0; t < NT; ++t) part[blk * WSZ + (t * NC + c) * NC + cp] = acc[t];
}

// Sum partials: out[i] = 2 * sum_blk part[blk][i] (real gradient of Re/Im).
__global__ void k_reduce_cplx(const float2* __restrict__ part, int nblk, int width, float2* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  double ax = 0, ay = 0;
  for (int b = 0; b < nblk; ++b) {
    const float2 v = part[(size_t)b * width + i];
    ax += v.x;
    ay += v.y;
  }
  out[i] = make_float2((float)(2.0 * ax), (float)(2.0 * ay));
}
__global__ void k_reduce_real(const float* __restrict__ part, int nblk, int width, float* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  double a = 0;
  for (int b = 0; b < nblk; ++b) a += part[(size_t)b * width + i];
  out[i] = (float)a;
}

// Gaussian inputs from a counter hash of (seed, element).
template <class A>
__global__ void k_net_input(A* __restrict__ out, size_t total, unsigned long long seed, size_t offset) {
  const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= total) return;
  const size_t i = j + offset;
  auto mix = [](unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  };
  const unsigned long long h1 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * i + 1) * 0xD1B54A32D192ED03ull);
  const unsigned long long h2 = mix(seed * 0x9E3779B97F4A7C15ull + (2 * i + 2) * 0xD1B54A32D192ED03ull);
  const float u1 = ((float)(h1 >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float u2 = ((float)(h2 >> 40) + 0.5f) * (1.0f / 16777216.0f);
  const float rr = sqrtf(-2.f * logf(u1));
  act_io<A>::st(out, j, make_float2(rr * cosf(6.2831853f * u2), rr * sinf(6.2831853f * u2)));
}

__global__ void k_f32_to_act(const float2* __restrict__ in, bf16c* __restrict__ out, size_t total) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < total) act_io<bf16c>::st(out, i, in[i]);
}

template <class A>
__global__ void k_act_to_f64(const A* __restrict__ in, double2* __restrict__ out, size_t total) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= total) return;
  const float2 v = act_io<A>::ld(in, i);
  out[i] = make_double2(v.x, v.y);
}

// Column statistics of the loss: ny[ts][j] = |y_j|^2 (atomic row-chunk sums).
__global__ __launch_bounds__(256) void k_colnorm2(const double2* __restrict__ Y, double* __restrict__ ny, int n, int K,
                                                   size_t per_ts) {
  const int ts = blockIdx.z;
  const int col = blockIdx.y * 64 + (threadIdx.x & 63);
  const int r0 = blockIdx.x * 256 + (threadIdx.x >> 6), r1 = min(n, (int)(blockIdx.x + 1) * 256);
  double acc = 0;
  for (int r = r0; r < r1; r += 4) {
    const double2 v = Y[per_ts * ts + (size_t)r * K + col];
    acc += v.x * v.x + v.y * v.y;
  }
  __shared__ double part[256];
  part[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < 64)
    atomicAdd(&ny[(size_t)ts * K + col], part[threadIdx.x] + part[threadIdx.x + 64] + part[threadIdx.x + 128] +
                                             part[threadIdx.x + 192]);
}

// f[ts][j] = sum_i |St[ts][j + i*K]|^2 / ny[ts][j]
__global__ void k_fraction(const double2* __restrict__ St, const double* __restrict__ ny, double* __restrict__ f,
                           int K, int nev, int nts) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nts * K) return;
  const int ts = i / K, j = i % K;
  double acc = 0;
  for (int e = 0; e < nev; ++e) {
    const double2 v = St[(size_t)ts * K * nev + (size_t)e * K + j];
    acc += v.x * v.x + v.y * v.y;
  }
  f[i] = acc / ny[i];
}

// g = -(1/Kc) (PY - f Y) / |y|^2 in FP32
__global__ void k_loss_grad(const double2* __restrict__ Y, const double2* __restrict__ PY, const double* __restrict__ f,
                            const double* __restrict__ ny, float2* __restrict__ g, size_t per_ts, int K, double scale,
                            int nts) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= per_ts * nts) return;
  const size_t ts = i / per_ts;
  const int j = (int)(i % K);
  const double fj = f[ts * K + j], nj = ny[ts * K + j];
  const double2 y = Y[i], p = PY[i];
  g[i] = make_float2((float)(-scale * (p.x - fj * y.x) / nj), (float)(-scale * (p.y - fj * y.y) / nj));
}

}  // namespace laph

// ===========================================================================
struct laph_net {
  int layers = 4;
  int device = 0;
  std::vector<float> params;  // [layers][7][16][16][2] then [layers][16]
  std::vector<double> m, v;   // Adam moments
  long long step = 0;
  hipStream_t st = nullptr;
  rocblas_handle rb = nullptr;
  float2* dW = nullptr;  // device copy of the weights
  float* dB = nullptr;
  size_t cap = 0;        // capacity (bytes) of the scratch below
  void* scratch = nullptr;
};

namespace laph {

static size_t w_count(int layers) { return (size_t)layers * WSZ * 2; }
static size_t b_count(int layers) { return (size_t)layers * NC; }

static int net_upload(laph_net* nt) {
  LAPH_HIP(hipMemcpyAsync(nt->dW, nt->params.data(), w_count(nt->layers) * sizeof(float), hipMemcpyHostToDevice,
                          nt->st));
  LAPH_HIP(hipMemcpyAsync(nt->dB, nt->params.data() + w_count(nt->layers), b_count(nt->layers) * sizeof(float),
                          hipMemcpyHostToDevice, nt->st));
  return LAPH_OK;
}

static int grow(laph_net* nt, size_t bytes) {
  if (bytes <= nt->cap) return LAPH_OK;
  if (nt->scratch) LAPH_HIP(hipFree(nt->scratch));
  nt->scratch = nullptr;
  LAPH_HIP(hipMalloc(&nt->scratch, bytes));
  nt->cap = bytes;
  return LAPH_OK;
}

// Buffers of one forward/backward over nts timeslices of a [n][K] block.
template <class A>
struct NetBufs {
  std::vector<A*> act;   // layers+1
  std::vector<A*> z;     // layers-1 pre-activations
  float2 *g0, *g1;       // gradients
  double2 *Y64, *PY;     // loss
  double2* St;           // [nts][nev][K]
  double *ny, *f;
  float2* wpart;
  float* bpart;
  size_t nblk_w, nblk_b;
};

template <class A>
static int net_bufs(laph_net* nt, int L, int K, int nts, int nev, bool train, NetBufs<A>& B) {
  const int V3 = L * L * L, n = 3 * V3;
  const size_t per_ts = (size_t)n * K, tot = per_ts * nts;
  const size_t nbx = (V3 + 3) / 4, nwx = (V3 + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK;
  B.nblk_w = nwx * (K / 64) * nts;
  B.nblk_b = nbx * (K / 64) * nts;
  const int nl = nt->layers;
  const int nact = train ? nl + 1 : 2;
  size_t bytes = (size_t)nact * tot * sizeof(A);
  if (train)
    bytes += (size_t)(nl - 1) * tot * sizeof(A) + 2 * tot * sizeof(float2) + 2 * tot * sizeof(double2) +
             (size_t)nts * K * nev * sizeof(double2) + 2 * (size_t)nts * K * sizeof(double) +
             B.nblk_w * WSZ * sizeof(float2) + B.nblk_b * NC * sizeof(float) + 1024;
  LAPH_TRY(grow(nt, bytes));
  char* p = static_cast<char*>(nt->scratch);
  auto take = [&](size_t nb) {
    char* q = p;
    p += (nb + 255) / 256 * 256;
    return q;
  };
  B.act.assign(nl + 1, nullptr);
  for (int l = 0; l < nact; ++l) B.act[l] = reinterpret_cast<A*>(take(tot * sizeof(A)));
  if (!train) {  // inference: layer l reads act[l % 2] and writes act[(l + 1) % 2]
    for (int l = 2; l <= nl; ++l) B.act[l] = B.act[l % 2];
    return LAPH_OK;
  }
  B.z.assign(nl - 1, nullptr);
  for (int l = 0; l < nl - 1; ++l) B.z[l] = reinterpret_cast<A*>(take(tot * sizeof(A)));
  B.g0 = reinterpret_cast<float2*>(take(tot * sizeof(float2)));
  B.g1 = reinterpret_cast<float2*>(take(tot * sizeof(float2)));
  B.Y64 = reinterpret_cast<double2*>(take(tot * sizeof(double2)));
  B.PY = reinterpret_cast<double2*>(take(tot * sizeof(double2)));
  B.St = reinterpret_cast<double2*>(take((size_t)nts * K * nev * sizeof(double2)));
  B.ny = reinterpret_cast<double*>(take((size_t)nts * K * sizeof(double)));
  B.f = reinterpret_cast<double*>(take((size_t)nts * K * sizeof(double)));
  B.wpart = reinterpret_cast<float2*>(take(B.nblk_w * WSZ * sizeof(float2)));
  B.bpart = reinterpret_cast<float*>(take(B.nblk_b * NC * sizeof(float)));
  return LAPH_OK;
}

// Forward pass from act[0] (filled by the caller) through all layers.
template <class A>
static int net_forward(laph_net* nt, const float2* U32, int L, int K, int nts, NetBufs<A>& B, bool keep_z) {
  const int V3 = L * L * L;
  const size_t per_ts = (size_t)3 * V3 * K;
  dim3 grid((V3 + 3) / 4, K / 64, nts);
  for (int l = 0; l < nt->layers; ++l) {
    const bool last = l == nt->layers - 1;
    const float* bias = last ? nullptr : nt->dB + l * NC;
    A* zout = (!last && keep_z) ? B.z[l] : nullptr;
    k_net_mix<A, A, false><<<grid, 256, 0, nt->st>>>(nt->dW + (size_t)l * WSZ, bias, B.act[l], B.act[l + 1], zout, U32, L,
                                                     K, per_ts);
    LAPH_HIP(hipGetLastError());
  }
  return LAPH_OK;
}

// Loss and gradient w.r.t. the output (act[layers]) given targets V (the
// first nev columns of Xv, [n][K] FP64 layout, leading dimension Kv).
template <class A>
static int net_loss(laph_net* nt, int n, int K, int nts, const double2* Xv, int Kv, size_t per_ts_v, int nev,
                    NetBufs<A>& B, double* loss) {
  const size_t per_ts = (size_t)n * K, tot = per_ts * nts;
  k_act_to_f64<A><<<(unsigned)((tot + 255) / 256), 256, 0, nt->st>>>(B.act[nt->layers], B.Y64, tot);
  LAPH_HIP(hipGetLastError());
  using cplx = rocblas_double_complex;
  const cplx one(1.0, 0.0), zero(0.0, 0.0);
#define LAPH_RB_NET LAPH_RB
  // St (K x nev) = Yc Vc^H ; PYc (K x n) = St Vc
  LAPH_RB_NET(rocblas_zgemm_strided_batched(nt->rb, rocblas_operation_none, rocblas_operation_conjugate_transpose, K,
                                            nev, n, &one, reinterpret_cast<const cplx*>(B.Y64), K,
                                            (rocblas_stride)per_ts, reinterpret_cast<const cplx*>(Xv), Kv,
                                            (rocblas_stride)per_ts_v, &zero, reinterpret_cast<cplx*>(B.St), K,
                                            (rocblas_stride)K * nev, nts));
  LAPH_RB_NET(rocblas_zgemm_strided_batched(nt->rb, rocblas_operation_none, rocblas_operation_none, K, n, nev, &one,
                                            reinterpret_cast<const cplx*>(B.St), K, (rocblas_stride)K * nev,
                                            reinterpret_cast<const cplx*>(Xv), Kv, (rocblas_stride)per_ts_v, &zero,
                                            reinterpret_cast<cplx*>(B.PY), K, (rocblas_stride)per_ts, nts));
  LAPH_HIP(hipMemsetAsync(B.ny, 0, (size_t)nts * K * sizeof(double), nt->st));
  dim3 gn((n + 255) / 256, K / 64, nts);
  k_colnorm2<<<gn, 256, 0, nt->st>>>(B.Y64, B.ny, n, K, per_ts);
  k_fraction<<<(nts * K + 255) / 256, 256, 0, nt->st>>>(B.St, B.ny, B.f, K, nev, nts);
  const double scale = 1.0 / ((double)K * nts);
  k_loss_grad<<<(unsigned)((tot + 255) / 256), 256, 0, nt->st>>>(B.Y64, B.PY, B.f, B.ny, B.g0, per_ts, K, scale, nts);
  LAPH_HIP(hipGetLastError());
  std::vector<double> f((size_t)nts * K);
  LAPH_HIP(hipMemcpyAsync(f.data(), B.f, f.size() * sizeof(double), hipMemcpyDeviceToHost, nt->st));
  LAPH_HIP(hipStreamSynchronize(nt->st));
  double s = 0;
  for (double v : f) s += v;
  *loss = -s * scale;
  return LAPH_OK;
}

// Backward from g0 = dL/d(output); fills grad (params layout, real).
template <class A>
static int net_backward(laph_net* nt, const float2* U32, int L, int K, int nts, NetBufs<A>& B, std::vector<float>& grad) {
  const int V3 = L * L * L;
  const size_t per_ts = (size_t)3 * V3 * K;
  const int nl = nt->layers;
  grad.assign(nt->params.size(), 0.f);
  dim3 grid((V3 + 3) / 4, K / 64, nts);
  dim3 gw((V3 + SITES_PER_BLOCK - 1) / SITES_PER_BLOCK, K / 64, nts);
  float2* dWg = nullptr;
  float* dBg = nullptr;
  LAPH_HIP(hipMalloc(&dWg, (size_t)WSZ * sizeof(float2)));
  if (hipMalloc(&dBg, NC * sizeof(float)) != hipSuccess) {
    (void)hipFree(dWg);
    return set_error(LAPH_ERR_NOMEM, "net backward scratch");
  }
  int rc = LAPH_OK;
  float2* g = B.g0;   // gradient w.r.t. the layer output a_{l+1}
  float2* gz = B.g1;  // gradient w.r.t. the pre-activation z_{l+1}
  for (int l = nl - 1; l >= 0 && rc == LAPH_OK; --l) {
    const bool last = l == nl - 1;
    float2* gpre = g;
    if (!last) {
      k_net_gate_bwd<A><<<grid, 256, 0, nt->st>>>(B.z[l], nt->dB + l * NC, g, gz, B.bpart, V3, K, per_ts);
      k_reduce_real<<<1, 64, 0, nt->st>>>(B.bpart, (int)B.nblk_b, NC, dBg);
      std::vector<float> hb(NC);
      if (hipMemcpyAsync(hb.data(), dBg, NC * sizeof(float), hipMemcpyDeviceToHost, nt->st) != hipSuccess ||
          hipStreamSynchronize(nt->st) != hipSuccess) {
        rc = set_error(LAPH_ERR_HIP, "net backward: bias gradient copy");
        break;
      }
      for (int c = 0; c < NC; ++c) grad[w_count(nl) + (size_t)l * NC + c] = hb[c];
      gpre = gz;
    }
    // weight gradient of layer l: inputs act[l], output gradient gpre
    k_net_wgrad<A><<<gw, 256, 0, nt->st>>>(B.act[l], gpre, U32, B.wpart, L, K, per_ts);
    k_reduce_cplx<<<(WSZ + 255) / 256, 256, 0, nt->st>>>(B.wpart, (int)B.nblk_w, WSZ, dWg);
    std::vector<float2> hw(WSZ);
    if (hipMemcpyAsync(hw.data(), dWg, WSZ * sizeof(float2), hipMemcpyDeviceToHost, nt->st) != hipSuccess ||
        hipStreamSynchronize(nt->st) != hipSuccess) {
      rc = set_error(LAPH_ERR_HIP, "net backward: weight gradient copy");
      break;
    }
    std::memcpy(grad.data() + (size_t)l * WSZ * 2, hw.data(), WSZ * sizeof(float2));
    // input gradient (not needed below layer 0)
    if (l > 0) {
      float2* gin = (gpre == B.g0) ? B.g1 : B.g0;
      k_net_mix<float2, float2, true><<<grid, 256, 0, nt->st>>>(nt->dW + (size_t)l * WSZ, nullptr, gpre, gin, nullptr,
                                                                U32, L, K, per_ts);
      if (hipGetLastError() != hipSuccess) {
        rc = set_error(LAPH_ERR_HIP, "net backward: adjoint layer launch");
        break;
      }
      g = gin;
      gz = (gin == B.g0) ? B.g1 : B.g0;
    }
  }
  (void)hipFree(dWg);
  (void)hipFree(dBg);
  return rc;
}

static void adam(laph_net* nt, const std::vector<float>& grad, double lr) {
  const double b1 = 0.9, b2 = 0.999, eps = 1e-8;
  ++nt->step;
  const double c1 = 1.0 - std::pow(b1, (double)nt->step), c2 = 1.0 - std::pow(b2, (double)nt->step);
  for (size_t i = 0; i < nt->params.size(); ++i) {
    nt->m[i] = b1 * nt->m[i] + (1 - b1) * grad[i];
    nt->v[i] = b2 * nt->v[i] + (1 - b2) * (double)grad[i] * grad[i];
    nt->params[i] -= (float)(lr * (nt->m[i] / c1) / (std::sqrt(nt->v[i] / c2) + eps));
  }
}

}  // namespace laph

// ---------------------------------------------------------------------------
extern "C" {

int laph_net_create(laph_net** out, int layers, unsigned long long seed, int device) {
  LAPH_CHECK(out && layers >= 1 && layers <= 16, LAPH_ERR_INVALID, "laph_net_create: bad arguments");
  *out = nullptr;
  LAPH_HIP(hipSetDevice(device));
  return laph::guarded([&]() -> int {
    laph_net* nt = new laph_net();
    nt->layers = layers;
    nt->device = device;
    const size_t nw = laph::w_count(layers), nb = laph::b_count(layers);
    nt->params.assign(nw + nb, 0.f);
    // initial network: each layer (1/2)(1 + H/6) on every channel -- a
    // low-pass polynomial in the hopping term -- plus a small random mixing
    uint64_t st = seed ? seed : 1;
    auto u = [&st]() {
      st ^= st << 13;
      st ^= st >> 7;
      st ^= st << 17;
      return ((st >> 11) + 0.5) * (1.0 / 9007199254740992.0) - 0.5;
    };
    for (int l = 0; l < layers; ++l)
      for (int t = 0; t < laph::NT; ++t)
        for (int c = 0; c < laph::NC; ++c)
          for (int cp = 0; cp < laph::NC; ++cp) {
            float* w = &nt->params[((((size_t)l * laph::NT + t) * laph::NC + c) * laph::NC + cp) * 2];
            const double diag = (c == cp) ? (t == 0 ? 0.5 : 1.0 / 12.0) : 0.0;
            w[0] = (float)(diag + 0.02 * u());
            w[1] = (float)(0.02 * u());
          }
    nt->m.assign(nt->params.size(), 0.0);
    nt->v.assign(nt->params.size(), 0.0);
    auto fail = [&](int rc) {
      laph_net_destroy(nt);
      return rc;
    };
    if (hipStreamCreateWithFlags(&nt->st, hipStreamNonBlocking) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_net_create: stream"));
    if (rocblas_create_handle(&nt->rb) != rocblas_status_success || rocblas_set_stream(nt->rb, nt->st) != rocblas_status_success)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_net_create: rocBLAS handle"));
    if (hipMalloc(&nt->dW, nw * sizeof(float)) != hipSuccess || hipMalloc(&nt->dB, nb * sizeof(float)) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_NOMEM, "laph_net_create: weights"));
    int rc = laph::net_upload(nt);
    if (rc != LAPH_OK) return fail(rc);
    *out = nt;
    return LAPH_OK;
  });
}

int laph_net_destroy(laph_net* nt) {
  if (!nt) return LAPH_OK;
  (void)hipSetDevice(nt->device);
  if (nt->st) (void)hipStreamSynchronize(nt->st);
  if (nt->dW) (void)hipFree(nt->dW);
  if (nt->dB) (void)hipFree(nt->dB);
  if (nt->scratch) (void)hipFree(nt->scratch);
  if (nt->rb) rocblas_destroy_handle(nt->rb);
  if (nt->st) (void)hipStreamDestroy(nt->st);
  delete nt;
  return LAPH_OK;
}

int laph_net_num_params(laph_net* nt) { return nt ? (int)nt->params.size() : 0; }

int laph_net_get_params(laph_net* nt, float* p) {
  LAPH_CHECK(nt && p, LAPH_ERR_INVALID, "laph_net_get_params: null argument");
  std::memcpy(p, nt->params.data(), nt->params.size() * sizeof(float));
  return LAPH_OK;
}

int laph_net_set_params(laph_net* nt, const float* p) {
  LAPH_CHECK(nt && p, LAPH_ERR_INVALID, "laph_net_set_params: null argument");
  LAPH_HIP(hipSetDevice(nt->device));
  std::memcpy(nt->params.data(), p, nt->params.size() * sizeof(float));
  std::fill(nt->m.begin(), nt->m.end(), 0.0);
  std::fill(nt->v.begin(), nt->v.end(), 0.0);
  nt->step = 0;
  LAPH_TRY(laph::net_upload(nt));
  LAPH_HIP(hipStreamSynchronize(nt->st));
  return LAPH_OK;
}

int laph_net_loss_grad(laph_net* nt, int L, int K, int fp32_act, const double* links, const float* input,
                       const double* V, int nev, double* loss, double* output, float* grad) {
  LAPH_CHECK(nt && links && input && V && loss && L >= 3 && K >= 64 && K % 64 == 0 && nev >= 1, LAPH_ERR_INVALID,
             "laph_net_loss_grad: bad arguments");
  LAPH_HIP(hipSetDevice(nt->device));
  return laph::guarded([&]() -> int {
    const int V3 = L * L * L, n = 3 * V3;
    const size_t per = (size_t)n * K, lk = (size_t)V3 * 27;
    // device links (FP32), targets ([n][nev] FP64 as a K-wide block with Kv = nev)
    float2* dU = nullptr;
    double2* dV = nullptr;
    double2* dU64 = nullptr;
    LAPH_HIP(hipMalloc(&dU, lk * sizeof(float2)));
    auto cleanup = [&]() {
      if (dU) (void)hipFree(dU);
      if (dV) (void)hipFree(dV);
      if (dU64) (void)hipFree(dU64);
    };
    int rc = LAPH_OK;
    if (hipMalloc(&dV, (size_t)n * nev * sizeof(double2)) != hipSuccess ||
        hipMalloc(&dU64, lk * sizeof(double2)) != hipSuccess) {
      cleanup();
      return laph::set_error(LAPH_ERR_NOMEM, "laph_net_loss_grad: device buffers");
    }
    (void)hipMemcpyAsync(dU64, links, lk * sizeof(double2), hipMemcpyHostToDevice, nt->st);
    (void)laph::launch_convert_links32(dU64, dU, lk, nt->st);
    (void)hipMemcpyAsync(dV, V, (size_t)n * nev * sizeof(double2), hipMemcpyHostToDevice, nt->st);
    std::vector<float> g;
    auto run = [&](auto tag) -> int {
      using A = decltype(tag);
      laph::NetBufs<A> B;
      LAPH_TRY(laph::net_bufs<A>(nt, L, K, 1, nev, true, B));
      // input: host FP32 complex [n][K] -> activation storage
      float2* tmp = reinterpret_cast<float2*>(B.g1);
      LAPH_HIP(hipMemcpyAsync(tmp, input, per * sizeof(float2), hipMemcpyHostToDevice, nt->st));
      if (sizeof(A) == sizeof(float2)) {
        LAPH_HIP(hipMemcpyAsync(B.act[0], tmp, per * sizeof(float2), hipMemcpyDeviceToDevice, nt->st));
      } else {
        laph::k_f32_to_act<<<(unsigned)((per + 255) / 256), 256, 0, nt->st>>>(tmp, reinterpret_cast<laph::bf16c*>(B.act[0]), per);
      }
      LAPH_TRY(laph::net_forward<A>(nt, dU, L, K, 1, B, true));
      LAPH_TRY(laph::net_loss<A>(nt, n, K, 1, dV, nev, (size_t)n * nev, nev, B, loss));
      if (output) {
        LAPH_HIP(hipMemcpyAsync(output, B.Y64, per * sizeof(double2), hipMemcpyDeviceToHost, nt->st));
        LAPH_HIP(hipStreamSynchronize(nt->st));
      }
      LAPH_TRY(laph::net_backward<A>(nt, dU, L, K, 1, B, g));
      return LAPH_OK;
    };
    rc = fp32_act ? run(float2{}) : run(laph::bf16c{});
    cleanup();
    if (rc == LAPH_OK && grad) std::memcpy(grad, g.data(), g.size() * sizeof(float));
    return rc;
  });
}

int laph_net_propose(laph_net* nt, laph_batch* b, unsigned long long seed) {
  LAPH_CHECK(nt && b, LAPH_ERR_INVALID, "laph_net_propose: null argument");
  LAPH_CHECK(nt->device == b->p.device, LAPH_ERR_INVALID, "laph_net_propose: network and batch on different devices");
  LAPH_HIP(hipSetDevice(nt->device));
  return laph::guarded([&]() -> int {
    LAPH_HIP(hipStreamSynchronize(b->st));
    const int K = b->K, L = b->p.L, V3 = b->V3;
    const size_t per = b->per_ts;
    // bound the two BF16 activation buffers to ~8 GB per chunk of timeslices
    const int chunk = (int)std::max<size_t>(1, std::min<size_t>(b->nts, (size_t)8e9 / (2 * per * sizeof(laph::bf16c))));
    for (int t0 = 0; t0 < b->nts; t0 += chunk) {
      const int nc = std::min(chunk, b->nts - t0);
      laph::NetBufs<laph::bf16c> B;
      LAPH_TRY(laph::net_bufs<laph::bf16c>(nt, L, K, nc, b->nev, false, B));
      const size_t tot = per * nc;
      laph::k_net_input<laph::bf16c><<<(unsigned)((tot + 255) / 256), 256, 0, nt->st>>>(B.act[0], tot, seed, per * t0);
      LAPH_HIP(hipGetLastError());
      LAPH_TRY(laph::net_forward<laph::bf16c>(nt, b->U32 + (size_t)t0 * V3 * 27, L, K, nc, B, false));
      laph::k_act_to_f64<laph::bf16c><<<(unsigned)((tot + 255) / 256), 256, 0, nt->st>>>(
          B.act[nt->layers], b->B[0] + per * t0, tot);
      LAPH_HIP(hipGetLastError());
    }
    LAPH_HIP(hipStreamSynchronize(nt->st));
    b->have_start = true;
    return LAPH_OK;
  });
}

int laph_net_train_step(laph_net* nt, laph_batch* b, unsigned long long seed, double lr, double* loss) {
  LAPH_CHECK(nt && b && loss && lr > 0, LAPH_ERR_INVALID, "laph_net_train_step: bad arguments");
  LAPH_CHECK(b->solved, LAPH_ERR_INVALID, "laph_net_train_step: the batch has no solution to train on");
  LAPH_CHECK(nt->device == b->p.device, LAPH_ERR_INVALID, "laph_net_train_step: network and batch on different devices");
  LAPH_HIP(hipSetDevice(nt->device));
  return laph::guarded([&]() -> int {
    LAPH_HIP(hipStreamSynchronize(b->st));
    const int K = b->K, L = b->p.L, nts = b->nts;
    const size_t tot = b->per_ts * nts;
    const int tgt = b->net_target_all ? K : b->nev;  // target span: the nev eigenvectors or all K Ritz vectors
    laph::NetBufs<laph::bf16c> B;
    LAPH_TRY(laph::net_bufs<laph::bf16c>(nt, L, K, nts, tgt, true, B));
    laph::k_net_input<laph::bf16c><<<(unsigned)((tot + 255) / 256), 256, 0, nt->st>>>(B.act[0], tot, seed, 0);
    LAPH_HIP(hipGetLastError());
    LAPH_TRY(laph::net_forward<laph::bf16c>(nt, b->U32, L, K, nts, B, true));
    LAPH_TRY(laph::net_loss<laph::bf16c>(nt, b->n, K, nts, b->B[0], K, b->per_ts, tgt, B, loss));
    std::vector<float> grad;
    LAPH_TRY(laph::net_backward<laph::bf16c>(nt, b->U32, L, K, nts, B, grad));
    laph::adam(nt, grad, lr);
    LAPH_TRY(laph::net_upload(nt));
    LAPH_HIP(hipStreamSynchronize(nt->st));
    return LAPH_OK;
  });
}

}  // extern "C"
