// C-ABI entry points of the per-timeslice context: link upload
// (map_timeslice_to_eigen), HYP, and the MatShell callbacks.
#include <cstring>
#include <mutex>
#include <vector>

#include "ctx.h"

namespace laph {

// Recurrence buffers of laph_shell_mult_noalloc: two vectors of up to
// 3 * 48^3 complex doubles (10.6 MB), held by every device the library runs
// on until the process ends.  Host code never passes `scratch` itself to a
// kernel: its device address comes from hipGetSymbolAddress, per device.
constexpr int kNoallocMaxL = 48;
constexpr size_t kNoallocVec = 3ull * kNoallocMaxL * kNoallocMaxL * kNoallocMaxL;
static __device__ __align__(256) double2 scratch[2 * kNoallocVec];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double2* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on device `dev` (current), looked up on the
// first call on that device.
static int noalloc_scratch(int dev, double2** out) {
  LAPH_CHECK(dev >= 0 && dev < kMaxDevices, LAPH_ERR_INVALID, "device %d beyond the scratch table", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    LAPH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    LAPH_CHECK(p != nullptr, LAPH_ERR_HIP, "hipGetSymbolAddress returned no address on device %d", dev);
    g_scratch_ptr[dev] = static_cast<double2*>(p);
  }
  *out = g_scratch_ptr[dev];
  return LAPH_OK;
}

int ensure_packed(laph_ctx* ctx) {
  if (!ctx->links_dirty) return LAPH_OK;
  LAPH_HIP(launch_pack_links64(ctx->U, ctx->Us, ctx->p.L, ctx->stream));
  ctx->links_dirty = false;
  return LAPH_OK;
}

// T0 = x, T1 = B x, T_k = 2 B T_{k-1} - T_{k-2}  (shell_matop.cpp:150-166);
// the last step writes straight into y.
int cheb_apply(laph_ctx* ctx, const double2* x, double2* y) {
  LAPH_TRY(ensure_packed(ctx));
  const int L = ctx->p.L, deg = ctx->p.deg;
  hipStream_t st = ctx->stream;
  if (deg == 1) {
    LAPH_HIP(launch_cheb_step64(ctx->Us, x, nullptr, y, L, ctx->c, ctx->a, 1.0, 0.0, st));
    return LAPH_OK;
  }
  const double2* old = x;
  double2* cur = ctx->t[0];
  LAPH_HIP(launch_cheb_step64(ctx->Us, x, nullptr, cur, L, ctx->c, ctx->a, 1.0, 0.0, st));
  for (int k = 2; k <= deg; ++k) {
    double2* out;
    if (k == deg) {
      out = y;
    } else {
      out = ctx->t[0];
      for (int j = 0; j < 3; ++j)
        if (ctx->t[j] != cur && ctx->t[j] != old) { out = ctx->t[j]; break; }
    }
    LAPH_HIP(launch_cheb_step64(ctx->Us, cur, old, out, L, ctx->c, ctx->a, 2.0, -1.0, st));
    old = cur;
    cur = out;
  }
  return LAPH_OK;
}

int laplacian_apply(laph_ctx* ctx, const double2* x, double2* y) {
  LAPH_TRY(ensure_packed(ctx));
  LAPH_HIP(launch_cheb_step64(ctx->Us, x, nullptr, y, ctx->p.L, 1.0, 0.0, 1.0, 0.0, ctx->stream));
  return LAPH_OK;
}

static int upload_config(laph_ctx* ctx, const double* ts_config) {
  // timeslice.cpp:46-73: U[s][dir](a,b) = config[s*72 + (dir+1)*18 + 6a + 2b]
  std::vector<double> host((size_t)ctx->V3 * 27 * 2);
  for (int s = 0; s < ctx->V3; ++s)
    for (int dir = 0; dir < 3; ++dir)
      std::memcpy(&host[((size_t)s * 3 + dir) * 18], ts_config + (size_t)s * 72 + (dir + 1) * 18,
                  18 * sizeof(double));
  LAPH_HIP(hipMemcpyAsync(ctx->U, host.data(), host.size() * sizeof(double), hipMemcpyHostToDevice,
                          ctx->stream));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->links_dirty = true;
  return LAPH_OK;
}

__global__ void k_unit_links(double2* U, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int e = i % 9;
  U[i] = make_double2((e % 4 == 0) ? 1.0 : 0.0, 0.0);
}

static void free_ctx(laph_ctx* ctx) {
  if (!ctx) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipSetDevice(ctx->p.device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  void* bufs[] = {ctx->U, ctx->Us, ctx->dec, ctx->Unew, ctx->vx, ctx->vy, ctx->t[0], ctx->t[1], ctx->t[2]};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  (void)hipSetDevice(dev);
  delete ctx;
}

}  // namespace laph

extern "C" {

int laph_ctx_create(laph_ctx** out, const laph_params* p, const double* ts_config) {
  LAPH_CHECK(out && p, LAPH_ERR_INVALID, "laph_ctx_create: null argument");
  *out = nullptr;
  LAPH_CHECK(p->L >= 3 && p->L <= 256, LAPH_ERR_INVALID, "laph_ctx_create: L=%d out of range [3,256]", p->L);
  LAPH_CHECK(p->deg >= 1, LAPH_ERR_INVALID,
             "laph_ctx_create: DEG=%d < 1 (T_DEG(B) needs DEG >= 1)", p->deg);
  LAPH_CHECK(p->lambda_l > p->lambda_c, LAPH_ERR_INVALID,
             "laph_ctx_create: lambda_l=%g must exceed lambda_c=%g", p->lambda_l, p->lambda_c);
  int ndev = 0;
  LAPH_HIP(hipGetDeviceCount(&ndev));
  LAPH_CHECK(p->device >= 0 && p->device < ndev, LAPH_ERR_INVALID,
             "laph_ctx_create: device %d not present (%d devices)", p->device, ndev);
  return laph::guarded([&]() -> int {
    laph_ctx* ctx = new laph_ctx();
    ctx->p = *p;
    ctx->V3 = p->L * p->L * p->L;
    ctx->n = 3 * (size_t)ctx->V3;
    ctx->c = 2.0 / (p->lambda_l - p->lambda_c);
    ctx->a = 1.0 + ctx->c * p->lambda_c;
    auto fail = [&](int rc) {
      laph::free_ctx(ctx);
      return rc;
    };
    if (hipSetDevice(p->device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_HIP, "laph_ctx_create: stream creation failed"));
    const size_t lb = (size_t)ctx->V3 * 27 * sizeof(double2);
    const size_t vb = ctx->n * sizeof(double2);
    if (hipMalloc(&ctx->U, lb) != hipSuccess || hipMalloc(&ctx->Us, lb) != hipSuccess ||
        hipMalloc(&ctx->vx, vb) != hipSuccess || hipMalloc(&ctx->vy, vb) != hipSuccess ||
        hipMalloc(&ctx->t[0], vb) != hipSuccess || hipMalloc(&ctx->t[1], vb) != hipSuccess ||
        hipMalloc(&ctx->t[2], vb) != hipSuccess)
      return fail(laph::set_error(LAPH_ERR_NOMEM, "laph_ctx_create: device allocation failed"));
    int rc;
    if (ts_config) {
      rc = laph::upload_config(ctx, ts_config);
    } else {
      const int nl = ctx->V3 * 27;
      laph::k_unit_links<<<(nl + 255) / 256, 256, 0, ctx->stream>>>(ctx->U, nl);
      rc = hipGetLastError() == hipSuccess ? LAPH_OK : laph::set_error(LAPH_ERR_HIP, "unit links");
    }
    if (rc != LAPH_OK) return fail(rc);
    *out = ctx;
    return LAPH_OK;
  });
}

int laph_ctx_destroy(laph_ctx* ctx) {
  laph::free_ctx(ctx);
  return LAPH_OK;
}

int laph_map_timeslice(laph_ctx* ctx, const double* ts_config) {
  LAPH_CHECK(ctx && ts_config, LAPH_ERR_INVALID, "laph_map_timeslice: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  return laph::upload_config(ctx, ts_config);
}

int laph_set_links(laph_ctx* ctx, const double* links) {
  LAPH_CHECK(ctx && links, LAPH_ERR_INVALID, "laph_set_links: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  LAPH_HIP(hipMemcpyAsync(ctx->U, links, (size_t)ctx->V3 * 27 * sizeof(double2), hipMemcpyHostToDevice,
                          ctx->stream));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->links_dirty = true;
  return LAPH_OK;
}

int laph_get_links(laph_ctx* ctx, double* links) {
  LAPH_CHECK(ctx && links, LAPH_ERR_INVALID, "laph_get_links: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  LAPH_HIP(hipMemcpyAsync(links, ctx->U, (size_t)ctx->V3 * 27 * sizeof(double2), hipMemcpyDeviceToHost,
                          ctx->stream));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  return LAPH_OK;
}

int laph_hyp_smear(laph_ctx* ctx, double alpha_1, double alpha_2, int iter) {
  LAPH_CHECK(ctx && iter >= 0, LAPH_ERR_INVALID, "laph_hyp_smear: bad arguments");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  if (iter == 0) return LAPH_OK;
  if (!ctx->dec) LAPH_HIP(hipMalloc(&ctx->dec, (size_t)ctx->V3 * 54 * sizeof(double2)));
  if (!ctx->Unew) LAPH_HIP(hipMalloc(&ctx->Unew, (size_t)ctx->V3 * 27 * sizeof(double2)));
  for (int it = 0; it < iter; ++it) {
    LAPH_HIP(laph::launch_hyp_level1(ctx->U, ctx->dec, ctx->p.L, alpha_2, ctx->stream));
    LAPH_HIP(laph::launch_hyp_level2(ctx->U, ctx->dec, ctx->Unew, ctx->p.L, alpha_1, ctx->stream));
    std::swap(ctx->U, ctx->Unew);
  }
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  ctx->links_dirty = true;
  return LAPH_OK;
}

int laph_shell_mult(void* vctx, const double* x, double* y) {
  laph_ctx* ctx = static_cast<laph_ctx*>(vctx);
  LAPH_CHECK(ctx && x && y, LAPH_ERR_INVALID, "laph_shell_mult: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  const size_t vb = ctx->n * sizeof(double2);
  LAPH_HIP(hipMemcpyAsync(ctx->vx, x, vb, hipMemcpyHostToDevice, ctx->stream));
  LAPH_TRY(laph::cheb_apply(ctx, ctx->vx, ctx->vy));
  LAPH_HIP(hipMemcpyAsync(y, ctx->vy, vb, hipMemcpyDeviceToHost, ctx->stream));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  return LAPH_OK;
}

int laph_shell_get_diagonal(void* vctx, double* diag) {
  laph_ctx* ctx = static_cast<laph_ctx*>(vctx);
  LAPH_CHECK(ctx && diag, LAPH_ERR_INVALID, "laph_shell_get_diagonal: null argument");
  for (size_t i = 0; i < ctx->n; ++i) {
    diag[2 * i] = -6.0;  // shell_matop.cpp:253
    diag[2 * i + 1] = 0.0;
  }
  return LAPH_OK;
}

int laph_shell_mult_device(laph_ctx* ctx, const double* dx, double* dy) {
  LAPH_CHECK(ctx && dx && dy && dx != dy, LAPH_ERR_INVALID,
             "laph_shell_mult_device: null or aliased vectors");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  LAPH_TRY(laph::cheb_apply(ctx, reinterpret_cast<const double2*>(dx), reinterpret_cast<double2*>(dy)));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  return LAPH_OK;
}

int laph_shell_mult_noalloc(const laph_params* p, const double* d_links, const double* dx, double* dy,
                            void* stream) {
  LAPH_CHECK(p && d_links && dx && dy && dx != dy && (const double*)dy != d_links, LAPH_ERR_INVALID,
             "laph_shell_mult_noalloc: null or aliased arguments");
  LAPH_CHECK(p->L >= 3 && p->L <= laph::kNoallocMaxL, LAPH_ERR_INVALID,
             "laph_shell_mult_noalloc: L=%d outside [3, %d] (the fixed scratch holds two vectors of 3*%d^3)", p->L,
             laph::kNoallocMaxL, laph::kNoallocMaxL);
  LAPH_CHECK(p->deg >= 1, LAPH_ERR_INVALID, "laph_shell_mult_noalloc: DEG=%d < 1", p->deg);
  LAPH_CHECK(p->lambda_l > p->lambda_c, LAPH_ERR_INVALID,
             "laph_shell_mult_noalloc: lambda_l=%g must exceed lambda_c=%g", p->lambda_l, p->lambda_c);
  int ndev = 0;
  LAPH_HIP(hipGetDeviceCount(&ndev));
  LAPH_CHECK(p->device >= 0 && p->device < ndev, LAPH_ERR_INVALID, "laph_shell_mult_noalloc: no device %d",
             p->device);
  LAPH_HIP(hipSetDevice(p->device));
  double2* t = nullptr;
  LAPH_TRY(laph::noalloc_scratch(p->device, &t));
  const int L = p->L, deg = p->deg;
  const size_t n = 3ull * L * L * L;
  const double c = 2.0 / (p->lambda_l - p->lambda_c), a = 1.0 + c * p->lambda_c;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const double2* U = reinterpret_cast<const double2*>(d_links);
  const double2* x = reinterpret_cast<const double2*>(dx);
  double2* y = reinterpret_cast<double2*>(dy);
  // T_1 = B x, T_k = 2 B T_{k-1} - T_{k-2} (shell_matop.cpp:150-166) with the
  // outputs rotating over {y, t0, t1} so that step k writes buf[(deg-k) % 3]:
  // the last step writes y, and no step's output is its input or T_{k-2}.
  double2* buf[3] = {y, t, t + n};
  const double2* old = nullptr;
  const double2* cur = x;
  for (int k = 1; k <= deg; ++k) {
    double2* out = buf[(deg - k) % 3];
    if (k == 1)
      LAPH_HIP(laph::launch_cheb_step64_aos(U, cur, nullptr, out, L, c, a, 1.0, 0.0, st));
    else
      LAPH_HIP(laph::launch_cheb_step64_aos(U, cur, old, out, L, c, a, 2.0, -1.0, st));
    old = cur;
    cur = out;
  }
  return LAPH_OK;
}

int laph_laplacian(laph_ctx* ctx, const double* x, double* y) {
  LAPH_CHECK(ctx && x && y, LAPH_ERR_INVALID, "laph_laplacian: null argument");
  LAPH_HIP(hipSetDevice(ctx->p.device));
  const size_t vb = ctx->n * sizeof(double2);
  LAPH_HIP(hipMemcpyAsync(ctx->vx, x, vb, hipMemcpyHostToDevice, ctx->stream));
  LAPH_TRY(laph::laplacian_apply(ctx, ctx->vx, ctx->vy));
  LAPH_HIP(hipMemcpyAsync(y, ctx->vy, vb, hipMemcpyDeviceToHost, ctx->stream));
  LAPH_HIP(hipStreamSynchronize(ctx->stream));
  return LAPH_OK;
}

}  // extern "C"
