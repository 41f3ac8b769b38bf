// Streaming-bandwidth ceiling of the device for the access mixes of the
// filter kernels: 3 reads + 1 write (a Chebyshev step with source and old
// terms), 2 reads + 1 write, 1 read + 1 write (copy) and 1 read, float4 per
// lane, grid-stride loops over 2 GiB arrays, plain and non-temporal accesses,
// several grid sizes.  Prints one JSON object: GB/s (bytes moved / time) of
// the best grid for each mix.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#define CK(x)                                                                            \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) {                                                              \
      std::fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));                \
      return 1;                                                                          \
    }                                                                                    \
  } while (0)

typedef float f4 __attribute__((ext_vector_type(4)));

template <bool NT>
__device__ __forceinline__ f4 ld(const f4* p) {
  return NT ? __builtin_nontemporal_load(p) : *p;
}
template <bool NT>
__device__ __forceinline__ void st(f4* p, f4 v) {
  if (NT)
    __builtin_nontemporal_store(v, p);
  else
    *p = v;
}

// NR reads (NR = 1..3) and one write per element; NR = 0 is read-only.
template <int NR, bool NT>
__global__ __launch_bounds__(256) void k_mix(const f4* __restrict__ a, const f4* __restrict__ b,
                                             const f4* __restrict__ c, f4* __restrict__ o, size_t n) {
  f4 acc = {0.f, 0.f, 0.f, 0.f};
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    if (NR == 0) {
      acc += ld<NT>(a + i);
      continue;
    }
    f4 v = ld<NT>(a + i);
    if (NR >= 2) v = v * 1.5f + ld<NT>(b + i);
    if (NR >= 3) v = v + 0.5f * ld<NT>(c + i);
    st<NT>(o + i, v);
  }
  if (NR == 0 && acc.x == 1234.5f) o[0] = acc;  // keeps the reads
}

template <int NR, bool NT>
static int run(const f4* a, const f4* b, const f4* c, f4* o, size_t n, int grid, hipEvent_t e0, hipEvent_t e1,
               float* ms) {
  for (int rep = 0; rep < 2; ++rep) {  // first round warms up
    CK(hipEventRecord(e0));
    for (int r = 0; r < 10; ++r) k_mix<NR, NT><<<grid, 256>>>(a, b, c, o, n);
    CK(hipGetLastError());
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(ms, e0, e1));
  }
  *ms /= 10.f;
  return 0;
}

int main() {
  const size_t bytes = (size_t)2 << 30;
  const size_t n = bytes / sizeof(f4);
  f4 *a, *b, *c, *o;
  CK(hipMalloc(&a, bytes));
  CK(hipMalloc(&b, bytes));
  CK(hipMalloc(&c, bytes));
  CK(hipMalloc(&o, bytes));
  CK(hipMemset(a, 0, bytes));
  CK(hipMemset(b, 0, bytes));
  CK(hipMemset(c, 0, bytes));
  CK(hipMemset(o, 0, bytes));
  hipEvent_t e0, e1;
  CK(hipEventCreate(&e0));
  CK(hipEventCreate(&e1));
  const int grids[] = {1024, 2048, 4096, 8192, 16384, 65536};
  const char* names[] = {"read", "1r1w", "2r1w", "3r1w"};
  std::string out = "{";
  for (int nr = 0; nr <= 3; ++nr)
    for (int nt = 0; nt < 2; ++nt) {
      double best = 0;
      int gbest = 0;
      for (int g : grids) {
        float ms = 0;
        int rc = 0;
        if (nr == 0) rc = nt ? run<0, true>(a, b, c, o, n, g, e0, e1, &ms) : run<0, false>(a, b, c, o, n, g, e0, e1, &ms);
        if (nr == 1) rc = nt ? run<1, true>(a, b, c, o, n, g, e0, e1, &ms) : run<1, false>(a, b, c, o, n, g, e0, e1, &ms);
        if (nr == 2) rc = nt ? run<2, true>(a, b, c, o, n, g, e0, e1, &ms) : run<2, false>(a, b, c, o, n, g, e0, e1, &ms);
        if (nr == 3) rc = nt ? run<3, true>(a, b, c, o, n, g, e0, e1, &ms) : run<3, false>(a, b, c, o, n, g, e0, e1, &ms);
        if (rc) return rc;
        const double moved = (double)bytes * (nr == 0 ? 1 : nr + 1);
        const double gbs = moved / (ms * 1e-3) / 1e9;
        if (gbs > best) {
          best = gbs;
          gbest = g;
        }
      }
      char buf[128];
      std::snprintf(buf, sizeof buf, "%s\"%s%s_GBps\": %.0f, \"%s%s_grid\": %d", out.size() > 1 ? ", " : "",
                    names[nr], nt ? "_nt" : "", best, names[nr], nt ? "_nt" : "", gbest);
      out += buf;
    }
  out += "}";
  std::printf("%s\n", out.c_str());
  return 0;
}
