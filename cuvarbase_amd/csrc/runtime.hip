// Context, errors, streams, events and memory of libgls.so (C ABI).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <mutex>
#include <vector>

#include "gls_internal.hpp"

// Workspace of gls_lomb_execute_static: a file-scope device array, one copy
// per GPU, held from the loading of this code object to the end of the
// process.  Host code never uses this name as an address: gls_init looks up
// the array's device address once per device (hipGetSymbolAddress) and
// launches take that address.
static_assert(gls::kScratchBytes == GLS_STATIC_SCRATCH_BYTES, "scratch size and header differ");
static __device__ __align__(256) char scratch[gls::kScratchBytes];

namespace gls {

namespace {
thread_local std::string g_err;
std::mutex g_mu;
bool g_init = false;
int g_device = -1;
hipStream_t g_aux[kAuxStreams] = {};

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per device

// Page-locked staging ring of the descriptor uploads: a byte ring with one
// event per in-flight span; a span is reused only after its copy completed.
constexpr size_t kStageBytes = size_t(16) << 20;
constexpr int kStageEvents = 256;
struct StageSpan {
    size_t off, len;
    int ev;
};
struct StageRing {
    char* base = nullptr;
    char* dev = nullptr;   // the ring's device address
    size_t head = 0;
    std::deque<StageSpan> inflight;
    hipEvent_t ev[kStageEvents] = {};
    std::vector<int> free_ev;
};
std::mutex g_stage_mu;
StageRing g_stage;

// with g_mu held and the device current
int stage_setup() {
    if (g_stage.base) return GLS_OK;
    void* p = nullptr;
    GLS_HIP_TRY(hipHostMalloc(&p, kStageBytes, hipHostMallocDefault));
    std::memset(p, 0, kStageBytes);   // fault the pages in now, not inside a copy
    {
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
            (void)hipGetLastError();
            d = nullptr;
        }
        g_stage.dev = (char*)d;
    }
    for (int i = 0; i < kStageEvents; ++i) {
        hipError_t e = hipEventCreateWithFlags(&g_stage.ev[i], hipEventDisableTiming);
        if (e != hipSuccess) {
            for (int j = 0; j < i; ++j) (void)hipEventDestroy(g_stage.ev[j]);
            (void)hipHostFree(p);
            return hip_fail(e, "hipEventCreateWithFlags (staging ring)");
        }
    }
    g_stage.free_ev.clear();
    for (int i = kStageEvents - 1; i >= 0; --i) g_stage.free_ev.push_back(i);
    g_stage.inflight.clear();
    g_stage.head = 0;
    g_stage.base = (char*)p;
    return GLS_OK;
}

void stage_teardown() {
    std::lock_guard<std::mutex> lk(g_stage_mu);
    if (!g_stage.base) return;
    for (auto& sp : g_stage.inflight) (void)hipEventSynchronize(g_stage.ev[sp.ev]);
    for (auto& e : g_stage.ev) (void)hipEventDestroy(e);
    (void)hipHostFree(g_stage.base);
    g_stage = StageRing{};
}
// Copies by a kernel (loads and stores over the bus from the shader cores,
// no DMA engine): the DMA path's hipMemcpyAsync has been measured holding the
// calling thread for 7-9 ms now and then in a served stream of small batches,
// with the device idle; a kernel copy is enqueued like any other launch.
// Grid-stride loops over 16-byte words (both ends 16-byte aligned), the rest
// byte by byte.
constexpr int kCopyThreads = 256;
__global__ __launch_bounds__(kCopyThreads) void copy16_kernel(const uint4* __restrict__ src,
                                                              uint4* __restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * kCopyThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kCopyThreads)
        dst[i] = src[i];
}
__global__ __launch_bounds__(kCopyThreads) void copy1_kernel(const unsigned char* __restrict__ src,
                                                             unsigned char* __restrict__ dst,
                                                             size_t n) {
    for (size_t i = (size_t)blockIdx.x * kCopyThreads + threadIdx.x; i < n;
         i += (size_t)gridDim.x * kCopyThreads)
        dst[i] = src[i];
}
// src and dst: device addresses (device memory, or page-locked host memory
// through its device pointer)
hipError_t kernel_copy(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    const bool al = ((uintptr_t)dst % 16 == 0) && ((uintptr_t)src % 16 == 0);
    const size_t n16 = al ? bytes / 16 : 0;
    if (n16 > 0) {
        const size_t blocks = std::min<size_t>((n16 + kCopyThreads - 1) / kCopyThreads, 2048);
        hipLaunchKernelGGL(copy16_kernel, dim3((unsigned)blocks), dim3(kCopyThreads), 0, s,
                           (const uint4*)src, (uint4*)dst, n16);
    }
    const size_t done = n16 * 16;
    if (bytes > done) {
        const size_t rem = bytes - done;
        const size_t blocks = std::min<size_t>((rem + kCopyThreads - 1) / kCopyThreads, 1024);
        hipLaunchKernelGGL(copy1_kernel, dim3((unsigned)blocks), dim3(kCopyThreads), 0, s,
                           (const unsigned char*)src + done, (unsigned char*)dst + done, rem);
    }
    return hipGetLastError();
}

// the device address of p: itself for device memory, the mapped address for
// page-locked host memory; nullptr for memory the device cannot reach
void* device_address(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (a.type == hipMemoryTypeDevice) return const_cast<void*>(p);
    if (a.type == hipMemoryTypeHost) {
        void* d = nullptr;
        if (hipHostGetDevicePointer(&d, const_cast<void*>(p), 0) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        return d;
    }
    return nullptr;
}

}  // namespace

hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return hipSuccess;
    std::lock_guard<std::mutex> lk(g_stage_mu);
    StageRing& R = g_stage;
    const size_t need = (bytes + 255) / 256 * 256;
    if (!R.base || need > kStageBytes / 2)
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    auto release = [&](std::deque<StageSpan>::iterator it) {
        R.free_ev.push_back(it->ev);
        return R.inflight.erase(it);
    };
    // spans whose copies have finished, oldest first
    while (!R.inflight.empty() && hipEventQuery(R.ev[R.inflight.front().ev]) == hipSuccess)
        release(R.inflight.begin());
    const size_t off = (R.head + need <= kStageBytes) ? R.head : 0;
    // wait for older copies still reading [off, off + need), and for a free event
    for (auto it = R.inflight.begin(); it != R.inflight.end();) {
        if (it->off < off + need && off < it->off + it->len) {
            hipError_t e = hipEventSynchronize(R.ev[it->ev]);
            if (e != hipSuccess) return e;
            it = release(it);
        } else {
            ++it;
        }
    }
    if (R.free_ev.empty()) {
        hipError_t e = hipEventSynchronize(R.ev[R.inflight.front().ev]);
        if (e != hipSuccess) return e;
        release(R.inflight.begin());
    }
    std::memcpy(R.base + off, src, bytes);
    // a kernel copy from the ring's mapped address (no DMA engine), or the
    // DMA path when the ring has no device address
    hipError_t e = R.dev ? kernel_copy(dst, R.dev + off, bytes, s)
                         : hipMemcpyAsync(dst, R.base + off, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    const int ev = R.free_ev.back();
    e = hipEventRecord(R.ev[ev], s);
    if (e != hipSuccess) return e;
    R.free_ev.pop_back();
    R.inflight.push_back({off, need, ev});
    R.head = off + need;
    return hipSuccess;
}

int static_scratch(void** ptr) {
    if (!g_init) return fail(GLS_E_NOT_INIT, "gls_init() has not been called");
    void* p = (g_device >= 0 && g_device < kMaxDevices) ? g_scratch[g_device] : nullptr;
    if (!p) return fail(GLS_E_NOT_INIT, "no static scratch address for this device");
    *ptr = p;
    return GLS_OK;
}

hipError_t aux_streams(hipStream_t* out) {
    std::lock_guard<std::mutex> lk(g_mu);
    for (int i = 0; i < kAuxStreams; ++i) {
        if (!g_aux[i]) {
            hipError_t e = hipStreamCreateWithFlags(&g_aux[i], hipStreamNonBlocking);
            if (e != hipSuccess) {
                g_aux[i] = nullptr;
                return e;
            }
        }
        out[i] = g_aux[i];
    }
    return hipSuccess;
}

void set_error(const std::string& msg) { g_err = msg; }

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what) {
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? GLS_E_OOM : GLS_E_HIP;
}

bool initialised() { return g_init; }

int require_init() {
    if (!g_init) return fail(GLS_E_NOT_INIT, "gls_init() has not been called");
    return GLS_OK;
}

}  // namespace gls

using namespace gls;

extern "C" {

const char* gls_last_error(void) { return g_err.c_str(); }

int gls_version(int* major, int* minor, int* patch) {
    if (major) *major = 0;
    if (minor) *minor = 1;
    if (patch) *patch = 0;
    return GLS_OK;
}

int gls_device_count(int* count) {
    if (!count) return fail(GLS_E_INVALID, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return GLS_OK;
}

int gls_init(int device) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_init && device == g_device) return GLS_OK;
    if (g_init && device != g_device)
        return fail(GLS_E_INVALID, "libgls is already initialised on another device");
    int n = 0;
    GLS_HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(GLS_E_INVALID, "device index out of range");
    GLS_HIP_TRY(hipSetDevice(device));
    if (device >= kMaxDevices) return fail(GLS_E_UNSUPPORTED, "device index above 63");
    // once per device, with the device current: the device address of this
    // GPU's copy of `scratch`
    if (!g_scratch[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!p) return fail(GLS_E_HIP, "hipGetSymbolAddress(scratch) returned NULL");
        g_scratch[device] = p;
    }
    {
        std::lock_guard<std::mutex> sl(g_stage_mu);
        int rc = stage_setup();
        if (rc) return rc;
    }
    g_device = device;
    g_init = true;
    return GLS_OK;
}

int gls_shutdown(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_init) return GLS_OK;
    (void)hipDeviceSynchronize();
    stage_teardown();
    for (auto& a : g_aux) {
        if (a) (void)hipStreamDestroy(a);
        a = nullptr;
    }
    g_init = false;
    g_device = -1;
    return GLS_OK;
}

int gls_device_info(char* name, int len, int* num_cus, size_t* total_mem) {
    int rc = require_init();
    if (rc) return rc;
    hipDeviceProp_t prop;
    GLS_HIP_TRY(hipGetDeviceProperties(&prop, g_device));
    if (name && len > 0) {
        std::snprintf(name, (size_t)len, "%s (%s)", prop.name, prop.gcnArchName);
    }
    if (num_cus) *num_cus = prop.multiProcessorCount;
    if (total_mem) *total_mem = prop.totalGlobalMem;
    return GLS_OK;
}

int gls_stream_create(void** stream) {
    int rc = require_init();
    if (rc) return rc;
    if (!stream) return fail(GLS_E_INVALID, "stream is NULL");
    hipStream_t s;
    GLS_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void*)s;
    return GLS_OK;
}

int gls_stream_destroy(void* stream) {
    if (!stream) return GLS_OK;
    GLS_HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return GLS_OK;
}

int gls_stream_synchronize(void* stream) {
    GLS_HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return GLS_OK;
}

int gls_stream_query(void* stream, int* done) {
    if (!done) return fail(GLS_E_INVALID, "done is NULL");
    hipError_t e = hipStreamQuery((hipStream_t)stream);
    if (e == hipSuccess) {
        *done = 1;
        return GLS_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return GLS_OK;
    }
    return hip_fail(e, "hipStreamQuery");
}

int gls_event_create(void** event) {
    int rc = require_init();
    if (rc) return rc;
    if (!event) return fail(GLS_E_INVALID, "event is NULL");
    hipEvent_t ev;
    GLS_HIP_TRY(hipEventCreate(&ev));
    *event = (void*)ev;
    return GLS_OK;
}

int gls_event_destroy(void* event) {
    if (!event) return GLS_OK;
    GLS_HIP_TRY(hipEventDestroy((hipEvent_t)event));
    return GLS_OK;
}

int gls_event_record(void* event, void* stream) {
    GLS_HIP_TRY(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return GLS_OK;
}

int gls_event_synchronize(void* event) {
    GLS_HIP_TRY(hipEventSynchronize((hipEvent_t)event));
    return GLS_OK;
}

int gls_event_query(void* event, int* done) {
    if (!done) return fail(GLS_E_INVALID, "done is NULL");
    hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) {
        *done = 1;
        return GLS_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return GLS_OK;
    }
    return hip_fail(e, "hipEventQuery");
}

int gls_event_elapsed_ms(void* start, void* stop, float* ms) {
    if (!ms) return fail(GLS_E_INVALID, "ms is NULL");
    GLS_HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return GLS_OK;
}

int gls_device_synchronize(void) {
    GLS_HIP_TRY(hipDeviceSynchronize());
    return GLS_OK;
}

int gls_malloc(void** ptr, size_t bytes) {
    int rc = require_init();
    if (rc) return rc;
    if (!ptr) return fail(GLS_E_INVALID, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipMalloc(ptr, bytes));
    return GLS_OK;
}

int gls_free(void* ptr) {
    if (!ptr) return GLS_OK;
    GLS_HIP_TRY(hipFree(ptr));
    return GLS_OK;
}

int gls_host_alloc(void** ptr, size_t bytes) {
    int rc = require_init();
    if (rc) return rc;
    if (!ptr) return fail(GLS_E_INVALID, "ptr is NULL");
    *ptr = nullptr;
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return GLS_OK;
}

int gls_host_free(void* ptr) {
    if (!ptr) return GLS_OK;
    GLS_HIP_TRY(hipHostFree(ptr));
    return GLS_OK;
}

int gls_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return GLS_OK;
}

int gls_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    return GLS_OK;
}

int gls_memcpy_kernel_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return GLS_OK;
    int rc = require_init();
    if (rc) return rc;
    if (!dst || !src) return fail(GLS_E_INVALID, "NULL pointer");
    void* d = device_address(dst);
    const void* sp = device_address(src);
    if (!d || !sp)
        return fail(GLS_E_INVALID, "gls_memcpy_kernel_async: both ends must be device memory "
                                   "or page-locked host memory (gls_host_alloc)");
    GLS_HIP_TRY(kernel_copy(d, sp, bytes, (hipStream_t)stream));
    return GLS_OK;
}

int gls_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return GLS_OK;
}

int gls_memset_async(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return GLS_OK;
    GLS_HIP_TRY(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return GLS_OK;
}

int gls_mem_info(size_t* free_bytes, size_t* total_bytes) {
    int rc = require_init();
    if (rc) return rc;
    size_t f = 0, t = 0;
    GLS_HIP_TRY(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return GLS_OK;
}

}  // extern "C"
