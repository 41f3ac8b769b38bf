// Issue rates of the arithmetic units the path's kernels use, with no memory
// traffic, to price instruction counts against what the device delivers:
//
//   fp32_fma    v_fma_f32, 8 independent chains per lane (the FP32 filter
//               step's link products; packed FP32 is not used)
//   fp64_fma    v_fma_f64, 8 independent chains per lane (the FP64 stencils)
//   fp64_mfma   v_mfma_f64_16x16x4f64, 4 independent accumulators per
//               wavefront (the Hermitian Gram products)
//
// Each runs 8 wavefronts per SIMD over the whole device.  The first lane of
// every work-group records shader-clock and constant-clock ticks over its
// loop, so the shader clock during the run is measured, not assumed.  Prints
// one JSON object: per unit the microseconds of the launch, TFLOP/s (2 flops
// per FMA), the measured clock in MHz and FMAs per cycle per SIMD.
// Usage: peak_rates.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <string>

#define CK(x)                                                                            \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                \
      std::exit(1);                                                                      \
    }                                                                                    \
  } while (0)

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void record(long long* clk, long long t0, long long w0) {
  const long long t1 = clock64(), w1 = wall_clock64();
  if (threadIdx.x == 0) {
    clk[2 * blockIdx.x] = t1 - t0;
    clk[2 * blockIdx.x + 1] = w1 - w0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_fma_rate(T* __restrict__ o, long long* __restrict__ clk, T b, T c,
                                                  int iters) {
  T a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = (T)(threadIdx.x + i) * (T)1e-3;
  const long long t0 = clock64(), w0 = wall_clock64();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = fma(a[i], b, c);
  }
  record(clk, t0, w0);
  T s = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += a[i];
  o[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_mfma64_rate(double* __restrict__ o, long long* __restrict__ clk, double a0,
                                                     double b0, int iters) {
  d4 acc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) acc[i] = d4{0.0, 0.0, 0.0, 0.0};
  const double a = a0 + threadIdx.x * 1e-6, b = b0;
  const long long t0 = clock64(), w0 = wall_clock64();
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  record(clk, t0, w0);
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) s += acc[i].x + acc[i].y + acc[i].z + acc[i].w;
  o[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

int main() {
  int dev = 0, ncu = 0, wall_khz = 0;
  CK(hipGetDevice(&dev));
  CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  CK(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev));
  const int nblk = ncu * 8;  // 4 wavefronts per work-group: 8 per SIMD
  double* out;
  long long* clk;
  CK(hipMalloc(&out, (size_t)nblk * 256 * sizeof(double)));
  CK(hipMalloc(&clk, (size_t)nblk * 2 * sizeof(long long)));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  std::string js = "{";
  char buf[320];
  std::snprintf(buf, sizeof buf, "\"compute_units\": %d, \"wall_clock_kHz\": %d", ncu, wall_khz);
  js += buf;
  // fmas = FMAs of one launch; the launch runs twice and the second is timed
  auto timeit = [&](const char* name, double fmas, auto launch) {
    float ms = 0;
    for (int round = 0; round < 2; ++round) {
      CK(hipEventRecord(e0));
      launch();
      CK(hipGetLastError());
      CK(hipEventRecord(e1));
      CK(hipEventSynchronize(e1));
      CK(hipEventElapsedTime(&ms, e0, e1));
    }
    long long c2[2] = {0, 0};
    CK(hipMemcpy(c2, clk, sizeof c2, hipMemcpyDeviceToHost));
    const double mhz = c2[1] > 0 ? (double)c2[0] / c2[1] * wall_khz * 1e-3 : 0.0;
    const double per_cycle = mhz > 0 ? fmas / (ms * 1e-3) / (ncu * 4.0 * mhz * 1e6) : 0.0;
    std::snprintf(buf, sizeof buf, ", \"%s_us\": %.1f, \"%s_TFLOPs\": %.2f, \"%s_clock_MHz\": %.0f, "
                  "\"%s_fma_per_cycle_per_simd\": %.2f", name, ms * 1e3, name, 2.0 * fmas / (ms * 1e-3) / 1e12, name,
                  mhz, name, per_cycle);
    js += buf;
  };
  const int it32 = 1 << 16, it64 = 1 << 15, itm = 1 << 13;
  timeit("fp32_fma", (double)nblk * 256 * 8 * it32, [&] {
    k_fma_rate<float><<<nblk, 256>>>(reinterpret_cast<float*>(out), clk, 0.999f, 1e-3f, it32);
  });
  timeit("fp64_fma", (double)nblk * 256 * 8 * it64, [&] {
    k_fma_rate<double><<<nblk, 256>>>(out, clk, 0.999, 1e-3, it64);
  });
  // one 16x16x4 MFMA = 1024 FMAs per wavefront, 4 wavefronts per work-group
  timeit("fp64_mfma", (double)nblk * 4 * 4 * 1024.0 * itm, [&] {
    k_mfma64_rate<<<nblk, 256>>>(out, clk, 1e-3, 1e-3, itm);
  });
  js += "}";
  std::printf("%s\n", js.c_str());
  CK(hipFree(out));
  CK(hipFree(clk));
  return 0;
}
