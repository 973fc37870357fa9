// Runtime part of the C ABI: handles, memory, copies, events, errors.
// This is the stand-in for the reference's sycl::queue + USM calls
// (main.cpp:189-192, Cholesky.cpp:24-32, CG.cpp:165-234).
#include <cstdio>
#include <cstring>
#include <string>

#include "hs_internal.h"

static thread_local std::string g_last_error;

void hs_set_error(const std::string& msg) { g_last_error = msg; }

int hs_hip_fail(hipError_t e, const char* what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return HS_ERR_HIP;
}

extern "C" const char* hs_last_error_string(void) { return g_last_error.c_str(); }

extern "C" int hs_device_count(int* count) {
    if (!count) return -1;
    HS_HIP(hipGetDeviceCount(count));
    return HS_OK;
}

extern "C" int hs_handle_create(hs_handle_t* handle, int device) {
    if (!handle) return -1;
    int count = 0;
    HS_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) {
        hs_set_error("hs_handle_create: no such device");
        return -2;
    }
    HS_HIP(hipSetDevice(device));
    hs_handle_s* h = new hs_handle_s();
    h->device = device;
    hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (h->stream) hipStreamDestroy(h->stream);
        delete h;
        return hs_hip_fail(e, "hipStreamCreateWithFlags");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) h->cu_count = prop.multiProcessorCount;
    e = hipMalloc((void**)&h->info_dev, sizeof(int) * 2);
    if (e == hipSuccess) e = hipMemset(h->info_dev, 0, sizeof(int) * 2);
    if (e == hipSuccess) e = hipHostMalloc(&h->pinned, 4096, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void**)&h->trsm_w, sizeof(double) * 16 * 256);
    if (e == hipSuccess) e = hipMalloc((void**)&h->hilo_scale, sizeof(float) * 64);
    for (int i = 0; i < HS_HANDLE_EVENTS && e == hipSuccess; ++i) {
        hipEvent_t ev;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e == hipSuccess) h->events.push_back(ev);
    }
    // the device address of the allocation-free CG's scratch, once per device
    if (e == hipSuccess && hs_noalloc_setup(device) != HS_OK) e = hipErrorInvalidSymbol;
    if (e != hipSuccess) {
        for (auto ev : h->events) hipEventDestroy(ev);
        if (h->hilo_scale) hipFree(h->hilo_scale);
        if (h->trsm_w) hipFree(h->trsm_w);
        if (h->pinned) hipHostFree(h->pinned);
        if (h->info_dev) hipFree(h->info_dev);
        hipStreamDestroy(h->stream2);
        hipStreamDestroy(h->stream);
        delete h;
        return hs_hip_fail(e, "hs_handle_create allocations");
    }
    *handle = h;
    return HS_OK;
}

extern "C" int hs_handle_destroy(hs_handle_t h) {
    if (!h) return -1;
    hipSetDevice(h->device);
    hipStreamSynchronize(h->stream);
    hipStreamSynchronize(h->stream2);
    for (auto e : h->events) hipEventDestroy(e);
    hipStreamDestroy(h->stream2);
    if (h->work) hipFree(h->work);
    if (h->info_dev) hipFree(h->info_dev);
    if (h->trsm_w) hipFree(h->trsm_w);
    if (h->hilo_scale) hipFree(h->hilo_scale);
    if (h->pinned) hipHostFree(h->pinned);
    hipStreamDestroy(h->stream);
    delete h;
    return HS_OK;
}

extern "C" int hs_sync(hs_handle_t h) {
    if (!h) return -1;
    HS_HIP(hipStreamSynchronize(h->stream));
    return HS_OK;
}

int hs_work(hs_handle_t h, size_t bytes, void** ptr) {
    if (bytes > h->work_bytes) {
        hs_set_error("handle workspace too small: " + std::to_string(bytes) + " bytes needed, " +
                     std::to_string(h->work_bytes) +
                     " reserved; reserve it at set-up with hs_workspace_reserve(handle, N, b, M)");
        return HS_ERR_WORKSPACE;
    }
    *ptr = h->work;
    return HS_OK;
}

int hs_ensure_events(hs_handle_t h, size_t count) {
    if (h->events.size() >= count) return HS_OK;
    hs_set_error("handle event pool too small: " + std::to_string(count) + " events needed, " +
                 std::to_string(h->events.size()) + " created; reserve at set-up with hs_workspace_reserve");
    return HS_ERR_WORKSPACE;
}

static int grow_events(hs_handle_t h, size_t count) {
    while (h->events.size() < count) {
        hipEvent_t e;
        HS_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        h->events.push_back(e);
    }
    return HS_OK;
}

// The largest need of any entry for a system of order N in blocks of b with
// M test points. The dot products over block ranges are bounded by the whole
// vector; the reference-order dot is taken at its smallest work-group (1).
static size_t workspace_need(i64 N, int b, i64 M) {
    size_t need = hs_ws_dot();
    auto take = [&](size_t v) { need = v > need ? v : need; };
    take(hs_ws_symv(N, b));
    take(hs_ws_cg(N, b));
    take(hs_ws_cg_ref(N, b, 1));
    take(hs_ws_dot_ref((N + b - 1) / b * b, 1));
    take(hs_ws_tri_solve(N, b));
    take(hs_ws_refine(N, b));
    take(hs_ws_mixed(N, b));
    take(hs_ws_adjoint(N, b));
    take(hs_ws_rbf_window(N));
    if (M > 0) take(hs_ws_gp_predict(N, M));
    return need;
}

extern "C" int hs_workspace_query(int64_t N, int b, int64_t M, size_t* bytes) {
    if (N <= 0) return -1;
    if (b <= 0) return -2;
    if (M < 0) return -3;
    if (!bytes) return -4;
    *bytes = workspace_need(N, b, M);
    return HS_OK;
}

extern "C" int hs_workspace_reserve(hs_handle_t h, int64_t N, int b, int64_t M) {
    if (!h) return -1;
    if (N <= 0) return -2;
    if (b <= 0) return -3;
    if (M < 0) return -4;
    HS_HIP(hipSetDevice(h->device));
    // the look-ahead factorisation orders its two streams with 2 n + 2 events
    int rc = grow_events(h, 2 * (size_t)((N + b - 1) / b) + 2);
    if (rc) return rc;
    const size_t need = workspace_need(N, b, M);
    if (need <= h->work_bytes) return HS_OK;
    // a set-up call: waits for the handle's work before the old buffer goes
    HS_HIP(hipStreamSynchronize(h->stream));
    HS_HIP(hipStreamSynchronize(h->stream2));
    if (h->work) HS_HIP(hipFree(h->work));
    h->work = nullptr;
    h->work_bytes = 0;
    hipError_t e = hipMalloc(&h->work, need);
    if (e != hipSuccess) {
        h->work = nullptr;
        hs_hip_fail(e, "hs_workspace_reserve");
        return HS_ERR_NOMEM;
    }
    h->work_bytes = need;
    return HS_OK;
}

extern "C" int hs_workspace_size(hs_handle_t h, size_t* bytes) {
    if (!h) return -1;
    if (!bytes) return -2;
    *bytes = h->work_bytes;
    return HS_OK;
}

extern "C" int hs_malloc(void** ptr, size_t bytes) {
    if (!ptr) return -1;
    hipError_t e = hipMalloc(ptr, bytes == 0 ? 1 : bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        hs_hip_fail(e, "hipMalloc");
        return HS_ERR_NOMEM;
    }
    return HS_OK;
}

extern "C" int hs_free(void* ptr) {
    if (ptr) HS_HIP(hipFree(ptr));
    return HS_OK;
}

extern "C" int hs_host_malloc(void** ptr, size_t bytes) {
    if (!ptr) return -1;
    hipError_t e = hipHostMalloc(ptr, bytes == 0 ? 1 : bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        hs_hip_fail(e, "hipHostMalloc");
        return HS_ERR_NOMEM;
    }
    return HS_OK;
}

extern "C" int hs_host_free(void* ptr) {
    if (ptr) HS_HIP(hipHostFree(ptr));
    return HS_OK;
}

extern "C" int hs_device_memory(hs_handle_t h, size_t* free_bytes, size_t* total_bytes) {
    if (!h) return -1;
    if (!free_bytes) return -2;
    if (!total_bytes) return -3;
    HS_HIP(hipSetDevice(h->device));
    HS_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return HS_OK;
}

extern "C" int hs_memcpy_htod(hs_handle_t h, void* dst, const void* src, size_t bytes) {
    if (!h) return -1;
    HS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    return HS_OK;
}

extern "C" int hs_memcpy_dtoh(hs_handle_t h, void* dst, const void* src, size_t bytes) {
    if (!h) return -1;
    HS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    return HS_OK;
}

extern "C" int hs_memcpy_dtod(hs_handle_t h, void* dst, const void* src, size_t bytes) {
    if (!h) return -1;
    HS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, h->stream));
    return HS_OK;
}

__global__ void hs_zero_kernel(double* p, i64 n64, unsigned char* tail, int ntail) {
    i64 i = (i64)blockIdx.x * blockDim.x + threadIdx.x;
    i64 stride = (i64)gridDim.x * blockDim.x;
    for (; i < n64; i += stride) p[i] = 0.0;
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) tail[threadIdx.x] = 0;
}

// A fill kernel rather than hipMemsetAsync, so callers may capture it in a graph.
extern "C" int hs_memset_zero(hs_handle_t h, void* dst, size_t bytes) {
    if (!h) return -1;
    if (bytes == 0) return HS_OK;
    if (((uintptr_t)dst & 7) != 0) {
        hs_set_error("hs_memset_zero: destination must be 8-byte aligned");
        return -2;
    }
    i64 n64 = (i64)(bytes / 8);
    int ntail = (int)(bytes % 8);
    i64 blocks = (n64 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    hs_zero_kernel<<<dim3((unsigned)blocks), dim3(256), 0, h->stream>>>((double*)dst, n64,
                                                                         (unsigned char*)dst + n64 * 8, ntail);
    HS_LAUNCH_CHECK();
    return HS_OK;
}

extern "C" int hs_event_create(hs_event_t* ev) {
    if (!ev) return -1;
    hs_event_s* e = new hs_event_s();
    hipError_t r = hipEventCreate(&e->ev);
    if (r != hipSuccess) {
        delete e;
        return hs_hip_fail(r, "hipEventCreate");
    }
    *ev = e;
    return HS_OK;
}

extern "C" int hs_event_destroy(hs_event_t ev) {
    if (!ev) return -1;
    hipEventDestroy(ev->ev);
    delete ev;
    return HS_OK;
}

extern "C" int hs_event_record(hs_handle_t h, hs_event_t ev) {
    if (!h) return -1;
    if (!ev) return -2;
    HS_HIP(hipEventRecord(ev->ev, h->stream));
    return HS_OK;
}

extern "C" int hs_event_elapsed_ms(hs_event_t start, hs_event_t stop, float* ms) {
    if (!start) return -1;
    if (!stop) return -2;
    if (!ms) return -3;
    HS_HIP(hipEventSynchronize(stop->ev));
    HS_HIP(hipEventElapsedTime(ms, start->ev, stop->ev));
    return HS_OK;
}

extern "C" int hs_get_info(hs_handle_t h, int64_t* info) {
    if (!h) return -1;
    if (!info) return -2;
    int v = 0;
    HS_HIP(hipMemcpyAsync(&v, h->info_dev, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HS_HIP(hipStreamSynchronize(h->stream));
    *info = v;
    return HS_OK;
}

extern "C" int hs_reset_info(hs_handle_t h) {
    if (!h) return -1;
    return hs_memset_zero(h, h->info_dev, 8);
}
