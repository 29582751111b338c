// wtg_ctx.hip -- context, memory, timing and fill kernels of the C ABI.
#include "wtg_internal.hpp"
#include <cstring>
#include <cstdlib>

static thread_local std::string g_err;
void wtg_set_error(const std::string& msg) { g_err = msg; }

extern "C" const char* wtg_version(void) { return "wtg 0.1 gfx950"; }
extern "C" const char* wtg_last_error(void) { return g_err.c_str(); }

extern "C" int wtg_device_count(int* out) {
  WTG_REQUIRE(out, "wtg_device_count: null out");
  int n = 0;
  WTG_CHECK_HIP(hipGetDeviceCount(&n));
  *out = n;
  return WTG_OK;
}

int wtg_set_device(wtg_ctx* ctx) {
  WTG_CHECK_HIP(hipSetDevice(ctx->device));
  return WTG_OK;
}

extern "C" int wtg_ctx_create(int device, wtg_ctx** out) {
  WTG_REQUIRE(out, "wtg_ctx_create: null out");
  int n = 0;
  WTG_CHECK_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    wtg_set_error("wtg_ctx_create: no such device");
    return WTG_ERR_ARG;
  }
  hipDeviceProp_t prop;
  WTG_CHECK_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    wtg_set_error(std::string("wtg_ctx_create: built for gfx950, device is ") + prop.gcnArchName);
    return WTG_ERR_UNSUPPORTED;
  }
  wtg_ctx* c = new wtg_ctx();
  c->device = device;
  WTG_CHECK_HIP(hipSetDevice(device));
  WTG_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  WTG_CHECK_HIP(hipMalloc(&c->d_flag, 256));
  WTG_CHECK_HIP(hipMemset(c->d_flag, 0, 256));
  WTG_CHECK_HIP(hipEventCreate(&c->t0));
  WTG_CHECK_HIP(hipEventCreate(&c->t1));
  WTG_CHECK_HIP(hipEventCreateWithFlags(&c->h2d_ev, hipEventDisableTiming));
  WTG_CHECK_HIP(hipHostMalloc((void**)&c->h_stage, WTG_STAGE_BYTES, hipHostMallocDefault));
  // library allocations come stream-ordered from the device's default pool,
  // which keeps freed memory for reuse (WTG_NO_POOL=1: plain hipMalloc)
  if (!std::getenv("WTG_NO_POOL")) {
    int pools = 0;
    hipMemPool_t pool = nullptr;
    uint64_t keep = UINT64_MAX;
    if (hipDeviceGetAttribute(&pools, hipDeviceAttributeMemoryPoolsSupported, device) == hipSuccess && pools &&
        hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess &&
        hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess)
      c->pooled = true;
    (void)hipGetLastError();
  }
  WTG_CHECK_HIP(hipDeviceSynchronize());
  {
    const int rc = wtg_noalloc_init_device(device);
    if (rc != WTG_OK) { wtg_ctx_destroy(c); return rc; }
  }
  *out = c;
  return WTG_OK;
}

extern "C" int wtg_ctx_destroy(wtg_ctx* ctx) {
  if (!ctx) return WTG_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& p : ctx->kpending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  for (auto e : ctx->kpool) (void)hipEventDestroy(e);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->d_flag) (void)hipFree(ctx->d_flag);
  if (ctx->t0) (void)hipEventDestroy(ctx->t0);
  if (ctx->t1) (void)hipEventDestroy(ctx->t1);
  if (ctx->h2d_ev) (void)hipEventDestroy(ctx->h2d_ev);
  if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return WTG_OK;
}

extern "C" int wtg_ctx_sync(wtg_ctx* ctx) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_TRY(wtg_set_device(ctx));
  // the flag rides the stream: one async copy into pinned memory, one sync
  volatile uint32_t* hf = reinterpret_cast<volatile uint32_t*>(ctx->h_stage);
  WTG_CHECK_HIP(hipMemcpyAsync(ctx->h_stage, ctx->d_flag, 4, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  const uint32_t f = *hf;
  if (f) {
    WTG_CHECK_HIP(hipMemset(ctx->d_flag, 0, 4));
    if (f & WTG_FLAG_BADSYM) {
      wtg_set_error("input symbol >= 2^levels (or inconsistent with the tree); results invalid");
      return WTG_ERR_ARG;
    }
    if (f & WTG_FLAG_SA_RANKS) {
      wtg_set_error("device suffix ranks inconsistent; results invalid");
      return WTG_ERR_DEVICE;
    }
    if (f & WTG_FLAG_BAD_SA) {
      wtg_set_error("suffix array entry outside [0, n): not a suffix array of this text; BWT invalid");
      return WTG_ERR_ARG;
    }
    wtg_set_error("device-side bounded wait expired (look-back scan); results invalid");
    return WTG_ERR_DEVICE;
  }
  return WTG_OK;
}

int wtg_stage_d2h(wtg_ctx* ctx, void* dst, const void* d_src, size_t bytes) {
  if (bytes > WTG_STAGE_D2H_MAX) {
    WTG_CHECK_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return WTG_OK;
  }
  WTG_CHECK_HIP(hipMemcpyAsync(ctx->h_stage + WTG_STAGE_D2H, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(dst, ctx->h_stage + WTG_STAGE_D2H, bytes);
  return WTG_OK;
}

int wtg_stage_h2d(wtg_ctx* ctx, void* d_dst, const void* src, size_t bytes) {
  if (bytes > WTG_STAGE_H2D_MAX) {
    WTG_CHECK_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    return WTG_OK;
  }
  WTG_CHECK_HIP(hipEventSynchronize(ctx->h2d_ev));  // the previous staged copy has read the buffer
  std::memcpy(ctx->h_stage + WTG_STAGE_H2D, src, bytes);
  WTG_CHECK_HIP(hipMemcpyAsync(d_dst, ctx->h_stage + WTG_STAGE_H2D, bytes, hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipEventRecord(ctx->h2d_ev, ctx->stream));
  return WTG_OK;
}

void wtg_abandon(wtg_ctx* ctx) {
  (void)hipGetLastError();
  if (hipStreamSynchronize(ctx->stream) == hipSuccess) (void)hipMemset(ctx->d_flag, 0, 4);
}

int wtg_finish(wtg_ctx* ctx) {
  WTG_CHECK_HIP(hipGetLastError());
  if (ctx->blocking) return wtg_ctx_sync(ctx);
  return WTG_OK;
}

extern "C" int wtg_ctx_set_level_method(wtg_ctx* ctx, int method) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_REQUIRE(method == 0 || method == 1, "wtg_ctx_set_level_method: method must be 0 or 1");
  ctx->level_method = method;
  return WTG_OK;
}

extern "C" int wtg_ctx_set_blocking(wtg_ctx* ctx, int blocking) {
  WTG_REQUIRE(ctx, "null ctx");
  ctx->blocking = blocking ? 1 : 0;
  return WTG_OK;
}
extern "C" void* wtg_ctx_stream(wtg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int wtg_ws_reserve(wtg_ctx* ctx, size_t bytes) {
  if (bytes <= ctx->ws_bytes) return WTG_OK;
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  if (ctx->ws) { WTG_CHECK_HIP(hipFree(ctx->ws)); ctx->ws = nullptr; ctx->ws_bytes = 0; }
  size_t want = bytes + bytes / 8;
  hipError_t e = hipMalloc(&ctx->ws, want);
  if (e != hipSuccess) {
    ctx->ws = nullptr;
    (void)hipGetLastError();
    wtg_set_error("workspace allocation failed");
    return WTG_ERR_NOMEM;
  }
  ctx->ws_bytes = want;
  return WTG_OK;
}

extern "C" int wtg_malloc(wtg_ctx* ctx, void** dptr, uint64_t bytes) {
  WTG_REQUIRE(ctx && dptr, "wtg_malloc: null argument");
  WTG_TRY(wtg_set_device(ctx));
  hipError_t e = ctx->pooled ? hipMallocAsync(dptr, bytes ? bytes : 1, ctx->stream) : hipMalloc(dptr, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *dptr = nullptr;
    wtg_set_error("hipMalloc failed");
    return WTG_ERR_NOMEM;
  }
  return WTG_OK;
}
extern "C" int wtg_free(wtg_ctx* ctx, void* dptr) {
  WTG_REQUIRE(ctx, "null ctx");
  if (!dptr) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  if (ctx->pooled) {
    WTG_CHECK_HIP(hipFreeAsync(dptr, ctx->stream));  // after the stream's earlier work
  } else {
    WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    WTG_CHECK_HIP(hipFree(dptr));
  }
  return WTG_OK;
}
extern "C" int wtg_memcpy_h2d(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WTG_REQUIRE(ctx && (bytes == 0 || (dst && src)), "wtg_memcpy_h2d: null argument");
  if (!bytes) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return WTG_OK;
}
extern "C" int wtg_memcpy_d2h(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WTG_REQUIRE(ctx && (bytes == 0 || (dst && src)), "wtg_memcpy_d2h: null argument");
  if (!bytes) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return WTG_OK;
}
extern "C" int wtg_memcpy_d2d(wtg_ctx* ctx, void* dst, const void* src, uint64_t bytes) {
  WTG_REQUIRE(ctx && (bytes == 0 || (dst && src)), "wtg_memcpy_d2d: null argument");
  if (!bytes) return WTG_OK;
  WTG_TRY(wtg_set_device(ctx));
  WTG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return wtg_finish(ctx);
}

// ------------------------------------------------------------ fill kernels
__global__ void k_fill_u64(uint64_t* __restrict__ d, uint64_t n, uint64_t v) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = v;
}
__global__ void k_fill_u8_zero(uint8_t* __restrict__ d, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = 0;
}

int wtg_fill_u64(wtg_ctx* ctx, void* d, uint64_t n, uint64_t v) {
  if (n == 0) return WTG_OK;
  uint64_t blocks = wtg_ceil_div(n, 256);
  if (blocks > 8192) blocks = 8192;
  WTG_LAUNCH(ctx, "fill", k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, (uint64_t*)d, n, v);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}
int wtg_fill_bytes_zero(wtg_ctx* ctx, void* d, uint64_t bytes) {
  if (bytes == 0) return WTG_OK;
  if (((uintptr_t)d & 7) == 0 && (bytes & 7) == 0) return wtg_fill_u64(ctx, d, bytes / 8, 0);
  uint64_t blocks = wtg_ceil_div(bytes, 256);
  if (blocks > 8192) blocks = 8192;
  WTG_LAUNCH(ctx, "fill", k_fill_u8_zero, dim3((unsigned)blocks), dim3(256), 0, (uint8_t*)d, bytes);
  WTG_LAUNCH_CHECK();
  return WTG_OK;
}
extern "C" int wtg_memset_zero(wtg_ctx* ctx, void* dst, uint64_t bytes) {
  WTG_REQUIRE(ctx && (bytes == 0 || dst), "wtg_memset_zero: null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(wtg_fill_bytes_zero(ctx, dst, bytes));
  return wtg_finish(ctx);
}

// ------------------------------------------------------------------ timing
extern "C" int wtg_timer_start(wtg_ctx* ctx) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_TRY(wtg_set_device(ctx));
  WTG_CHECK_HIP(hipEventRecord(ctx->t0, ctx->stream));
  return WTG_OK;
}
extern "C" int wtg_timer_stop(wtg_ctx* ctx, float* ms) {
  WTG_REQUIRE(ctx && ms, "null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_CHECK_HIP(hipEventRecord(ctx->t1, ctx->stream));
  WTG_CHECK_HIP(hipEventSynchronize(ctx->t1));
  WTG_CHECK_HIP(hipEventElapsedTime(ms, ctx->t0, ctx->t1));
  return WTG_OK;
}

static hipEvent_t kpool_get(wtg_ctx* ctx) {
  if (!ctx->kpool.empty()) { hipEvent_t e = ctx->kpool.back(); ctx->kpool.pop_back(); return e; }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}
wtg_kscope::wtg_kscope(wtg_ctx* c, const char* n) : ctx(c), name(n) {
  if (ctx->ktime && (ctx->kfilter.empty() || std::strncmp(n, ctx->kfilter.c_str(), ctx->kfilter.size()) == 0)) {
    a = kpool_get(ctx);
    b = kpool_get(ctx);
    (void)hipEventRecord(a, ctx->stream);
  }
}
wtg_kscope::~wtg_kscope() {
  if (a && b) {
    (void)hipEventRecord(b, ctx->stream);
    ctx->kpending.push_back({std::string(name), a, b});
  }
}
static int kflush(wtg_ctx* ctx) {
  if (ctx->kpending.empty()) return WTG_OK;
  WTG_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& p : ctx->kpending) {
    float ms = 0;
    WTG_CHECK_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    auto& s = ctx->kstats[p.name];
    s.ms += ms;
    s.launches += 1;
    ctx->kpool.push_back(p.a);
    ctx->kpool.push_back(p.b);
  }
  ctx->kpending.clear();
  return WTG_OK;
}
extern "C" int wtg_kernel_timing(wtg_ctx* ctx, int enable) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(kflush(ctx));
  ctx->ktime = enable != 0;
  return WTG_OK;
}

extern "C" int wtg_kernel_timing_only(wtg_ctx* ctx, const char* name_prefix) {
  WTG_REQUIRE(ctx, "null ctx");
  ctx->kfilter = name_prefix ? name_prefix : "";
  return WTG_OK;
}
extern "C" int wtg_kernel_time(wtg_ctx* ctx, const char* name, double* total_ms, uint64_t* launches) {
  WTG_REQUIRE(ctx && name && total_ms && launches, "null argument");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(kflush(ctx));
  auto it = ctx->kstats.find(name);
  *total_ms = (it == ctx->kstats.end()) ? 0.0 : it->second.ms;
  *launches = (it == ctx->kstats.end()) ? 0 : it->second.launches;
  return WTG_OK;
}
extern "C" int wtg_kernel_timing_reset(wtg_ctx* ctx) {
  WTG_REQUIRE(ctx, "null ctx");
  WTG_TRY(wtg_set_device(ctx));
  WTG_TRY(kflush(ctx));
  ctx->kstats.clear();
  return WTG_OK;
}
