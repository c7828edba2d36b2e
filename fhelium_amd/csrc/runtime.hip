// Device runtime of the C ABI: per-thread error text, version and device
// description (the gfx950 counterpart of the reference's CUDA inspection
// module csrc/runtime/cuda_info.cpp), plus the device services a host
// without PyTorch needs around the kernels:
//   * a streaming copy kernel (16-byte vector loads and stores), which is
//     also the byte reference for measured HBM ceilings and counter scaling;
//   * an arena: one device allocation carved into aligned workspaces by a
//     bump pointer, reset as a whole (fixed addresses, no allocator calls on
//     the hot path);
//   * timing events and stream-captured graphs (replay of a captured launch
//     sequence with one call).
#include <cstdarg>
#include <new>

#include "fh_common.h"

namespace fh {
namespace {
thread_local char g_error[512] = {0};
}  // namespace

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

const char* last_error() { return g_error; }

namespace {

// 16 bytes per thread per step; the grid covers the device a few times over
// and strides through the rest.
__global__ __launch_bounds__(256) void k_stream_copy(uint4* __restrict__ dst, const uint4* __restrict__ src,
                                                     int64_t vectors) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < vectors; i += stride) {
    dst[i] = src[i];
  }
}

__global__ void k_copy_tail(unsigned char* dst, const unsigned char* src, int64_t bytes) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < bytes) dst[i] = src[i];
}

int hip_fail(const char* what, hipError_t e) {
  set_error("%s: %s", what, hipGetErrorString(e));
  return kLaunchFailed;
}

}  // namespace
}  // namespace fh

using namespace fh;

FH_API const char* fh_last_error(void) { return fh::last_error(); }

FH_API int fh_abi_version(void) { return 7; }

// Writes "<arch> cu=<n> lds=<bytes> l2=<bytes> mem=<bytes>" for the current
// device into buf.
FH_API int fh_device_info(char* buf, int64_t buf_len) {
  FH_CHECK_ARG(buf && buf_len > 0, "device info: empty buffer");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail("hipGetDevice", e);
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return hip_fail("hipGetDeviceProperties", e);
  snprintf(buf, static_cast<size_t>(buf_len), "%s cu=%d lds=%zu l2=%d mem=%zu clock_khz=%d",
           p.gcnArchName, p.multiProcessorCount, p.sharedMemPerMultiprocessor,
           p.l2CacheSize, p.totalGlobalMem, p.clockRate);
  return kOk;
}

FH_API int fh_runtime_version(int* hip_runtime_version) {
  FH_CHECK_ARG(hip_runtime_version != nullptr, "runtime version: null output");
  const hipError_t e = hipRuntimeGetVersion(hip_runtime_version);
  if (e != hipSuccess) return hip_fail("hipRuntimeGetVersion", e);
  return kOk;
}

FH_API int fh_stream_copy(void* dst, const void* src, int64_t bytes, void* stream) {
  FH_CHECK_ARG(bytes >= 0, "stream copy: negative size");
  if (bytes == 0) return kOk;
  FH_CHECK_ARG(dst && src, "stream copy: null pointer");
  const auto d = reinterpret_cast<uintptr_t>(dst);
  const auto s = reinterpret_cast<uintptr_t>(src);
  FH_CHECK_ARG(d + static_cast<uintptr_t>(bytes) <= s || s + static_cast<uintptr_t>(bytes) <= d,
               "stream copy: source and destination overlap");
  FH_CHECK_ARG((d % 16) == 0 && (s % 16) == 0, "stream copy: pointers must be 16-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t vectors = bytes / 16;
  if (vectors > 0) {
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess) return hip_fail("stream copy: device query", e);
    const int64_t want = (vectors + 255) / 256;
    const int64_t cap = static_cast<int64_t>(cus) * 16;  // 16 blocks of 256 threads per CU
    const uint32_t blocks = static_cast<uint32_t>(want < cap ? want : cap);
    hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, st, static_cast<uint4*>(dst),
                       static_cast<const uint4*>(src), vectors);
    FH_CHECK_LAUNCH("stream copy");
  }
  const int64_t tail = bytes - vectors * 16;
  if (tail > 0) {
    hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(16), 0, st, static_cast<unsigned char*>(dst) + vectors * 16,
                       static_cast<const unsigned char*>(src) + vectors * 16, tail);
    FH_CHECK_LAUNCH("stream copy tail");
  }
  return kOk;
}

// ---- arena ---------------------------------------------------------------
struct fh_arena {
  unsigned char* base = nullptr;
  int64_t capacity = 0;
  int64_t used = 0;
  int64_t high_water = 0;
  int device = 0;
};

FH_API int fh_arena_create(int64_t bytes, fh_arena** out) {
  FH_CHECK_ARG(out != nullptr, "arena: null output");
  FH_CHECK_ARG(bytes > 0, "arena: capacity must be positive");
  *out = nullptr;
  fh_arena* a = new (std::nothrow) fh_arena();
  if (a == nullptr) {
    set_error("arena: host allocation failed");
    return kLaunchFailed;
  }
  hipError_t e = hipGetDevice(&a->device);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&a->base), static_cast<size_t>(bytes));
  if (e != hipSuccess) {
    delete a;
    return hip_fail("arena: hipMalloc", e);
  }
  a->capacity = bytes;
  *out = a;
  return kOk;
}

// Aligned bump allocation; fails (nothing changes) when the arena is full.
FH_API int fh_arena_alloc(fh_arena* arena, int64_t bytes, int64_t alignment, void** out) {
  FH_CHECK_ARG(arena != nullptr && out != nullptr, "arena: null argument");
  FH_CHECK_ARG(bytes >= 0, "arena: negative size");
  FH_CHECK_ARG(alignment > 0 && (alignment & (alignment - 1)) == 0 && alignment <= 4096,
               "arena: alignment must be a power of two up to 4096");
  const int64_t start = (arena->used + alignment - 1) & ~(alignment - 1);
  FH_CHECK_ARG(start <= arena->capacity && bytes <= arena->capacity - start,
               "arena: %lld bytes requested, %lld of %lld free", (long long)bytes,
               (long long)(arena->capacity - start < 0 ? 0 : arena->capacity - start),
               (long long)arena->capacity);
  *out = arena->base + start;
  arena->used = start + bytes;
  if (arena->used > arena->high_water) arena->high_water = arena->used;
  return kOk;
}

// Every earlier allocation becomes free; the caller orders this after the
// last use of them on the device.
FH_API int fh_arena_reset(fh_arena* arena) {
  FH_CHECK_ARG(arena != nullptr, "arena: null argument");
  arena->used = 0;
  return kOk;
}

FH_API int fh_arena_stats(const fh_arena* arena, int64_t* capacity, int64_t* used, int64_t* high_water) {
  FH_CHECK_ARG(arena != nullptr, "arena: null argument");
  if (capacity) *capacity = arena->capacity;
  if (used) *used = arena->used;
  if (high_water) *high_water = arena->high_water;
  return kOk;
}

FH_API int fh_arena_destroy(fh_arena* arena) {
  if (arena == nullptr) return kOk;
  int prev = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess && prev != arena->device) e = hipSetDevice(arena->device);
  if (e == hipSuccess) e = hipFree(arena->base);
  if (prev != arena->device) (void)hipSetDevice(prev);
  delete arena;
  if (e != hipSuccess) return hip_fail("arena: hipFree", e);
  return kOk;
}

// ---- events --------------------------------------------------------------
FH_API int fh_event_create(void** out) {
  FH_CHECK_ARG(out != nullptr, "event: null output");
  hipEvent_t ev = nullptr;
  const hipError_t e = hipEventCreate(&ev);
  if (e != hipSuccess) return hip_fail("hipEventCreate", e);
  *out = ev;
  return kOk;
}

FH_API int fh_event_record(void* event, void* stream) {
  FH_CHECK_ARG(event != nullptr, "event: null event");
  const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(event), static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail("hipEventRecord", e);
  return kOk;
}

// Milliseconds between two recorded events; waits for the second.
FH_API int fh_event_elapsed_ms(void* start, void* stop, float* ms) {
  FH_CHECK_ARG(start && stop && ms, "event: null argument");
  hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(stop));
  if (e == hipSuccess) e = hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop));
  if (e != hipSuccess) return hip_fail("hipEventElapsedTime", e);
  return kOk;
}

FH_API int fh_event_destroy(void* event) {
  if (event == nullptr) return kOk;
  const hipError_t e = hipEventDestroy(static_cast<hipEvent_t>(event));
  if (e != hipSuccess) return hip_fail("hipEventDestroy", e);
  return kOk;
}

// ---- stream-captured graphs ----------------------------------------------
// fh_graph_begin starts capturing every launch issued to `stream` (and to
// streams forked from it by events, as the fused key switch's side stream
// is); fh_graph_end stops and instantiates.  The ABI's entry points issue no
// memset or allocation, so a captured sequence holds kernel and event nodes
// only, and its workspaces keep the addresses they had at capture time.
FH_API int fh_graph_begin(void* stream) {
  const hipError_t e =
      hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail("hipStreamBeginCapture", e);
  return kOk;
}

FH_API int fh_graph_end(void* stream, void** graph_exec) {
  FH_CHECK_ARG(graph_exec != nullptr, "graph: null output");
  *graph_exec = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &g);
  if (e != hipSuccess) return hip_fail("hipStreamEndCapture", e);
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail("hipGraphInstantiate", e);
  *graph_exec = exec;
  return kOk;
}

FH_API int fh_graph_launch(void* graph_exec, void* stream) {
  FH_CHECK_ARG(graph_exec != nullptr, "graph: null graph");
  const hipError_t e = hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), static_cast<hipStream_t>(stream));
  if (e != hipSuccess) return hip_fail("hipGraphLaunch", e);
  return kOk;
}

FH_API int fh_graph_destroy(void* graph_exec) {
  if (graph_exec == nullptr) return kOk;
  const hipError_t e = hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec));
  if (e != hipSuccess) return hip_fail("hipGraphExecDestroy", e);
  return kOk;
}
