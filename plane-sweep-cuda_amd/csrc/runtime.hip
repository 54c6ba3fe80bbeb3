// runtime.hip -- status codes, error text, device memory, streams and events
// of the C ABI.  The reference reaches the CUDA runtime directly
// (src/inc/memory.h, src/inc/cuda_exception.h); FFI callers of libpsw get
// the same services through these entry points.
#include "psw_common.h"

#include <cstdarg>
#include <mutex>

namespace psw {

static thread_local char g_msg[512] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
    va_end(ap);
}

const char* get_error() { return g_msg; }

int hip_status(hipError_t e, const char* what, int code) {
    set_error("%s: %s (%s)", what, hipGetErrorString(e), hipGetErrorName(e));
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return PSW_ERR_ALLOC;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return PSW_ERR_DEVICE;
    return code;
}

int launch_status(const char* kernel) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_status(e, kernel, PSW_ERR_LAUNCH);
    return PSW_OK;
}

constexpr int WS_MAX_DEVICES = 64;

static int workspace_pool(hipMemPool_t* out) {
    static std::mutex mu;
    static hipMemPool_t pools[WS_MAX_DEVICES] = {};
    int dev = 0;
    PSW_TRY_HIP(hipGetDevice(&dev), PSW_ERR_DEVICE);
    if (dev < 0 || dev >= WS_MAX_DEVICES) {
        set_error("device index %d out of range", dev);
        return PSW_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(mu);
    if (!pools[dev]) {
        hipMemPoolProps props = {};
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        PSW_TRY_HIP(hipMemPoolCreate(&pool, &props), PSW_ERR_ALLOC);
        uint64_t keep = UINT64_MAX;   // never trim at synchronisations
        const hipError_t e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            return hip_status(e, "hipMemPoolSetAttribute(release threshold)", PSW_ERR_ALLOC);
        }
        pools[dev] = pool;
    }
    *out = pools[dev];
    return PSW_OK;
}

int workspace_alloc_async(void** p, size_t bytes, hipStream_t st) {
    hipMemPool_t pool = nullptr;
    const int rc = workspace_pool(&pool);
    if (rc != PSW_OK) return rc;
    PSW_TRY_HIP(hipMallocFromPoolAsync(p, bytes, pool, st), PSW_ERR_ALLOC);
    return PSW_OK;
}

int workspace_free_async(void* p, hipStream_t st) {
    if (p) PSW_TRY_HIP(hipFreeAsync(p, st), PSW_ERR_DEVICE);
    return PSW_OK;
}

int make_shape(psw_dim3 blocks, psw_dim3 threads, int width, int height, Shape& out) {
    if (width < 0 || height < 0) {
        set_error("negative image size %d x %d", width, height);
        return PSW_ERR_ARG;
    }
    const unsigned tcount = threads.x * threads.y * threads.z;
    const unsigned bcount = blocks.x * blocks.y * blocks.z;
    if (tcount == 0 || bcount == 0) {
        out.block = dim3(64, 4, 1);
        out.grid = dim3((unsigned)((width + 63) / 64), (unsigned)((height + 3) / 4), 1);
        if (out.grid.x == 0) out.grid.x = 1;
        if (out.grid.y == 0) out.grid.y = 1;
        return PSW_OK;
    }
    if (tcount > 1024) {
        set_error("block of %u threads exceeds the device maximum of 1024", tcount);
        return PSW_ERR_LAUNCH;
    }
    out.block = dim3(threads.x, threads.y, threads.z);
    out.grid = dim3(blocks.x, blocks.y, blocks.z);
    return PSW_OK;
}

__global__ void fill_f32_kernel(float* __restrict__ p, float v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void copy_f32_kernel(float4* __restrict__ dst, const float4* __restrict__ src, size_t n4) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

static unsigned fill_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b == 0) b = 1;
    return (unsigned)b;
}

}  // namespace psw

using namespace psw;

extern "C" {

const char* psw_status_string(int status) {
    switch (status) {
        case PSW_OK: return "ok";
        case PSW_ERR_ARG: return "invalid argument";
        case PSW_ERR_LAUNCH: return "kernel launch failed";
        case PSW_ERR_ALLOC: return "device allocation failed";
        case PSW_ERR_DEVICE: return "device error";
        case PSW_ERR_COPY: return "memory copy failed";
        case PSW_ERR_UNSUPPORTED: return "unsupported";
        default: return "unknown status";
    }
}

const char* psw_last_error(void) { return get_error(); }

int psw_device_count(int* count) {
    PSW_CHECK_ARG(count, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_status(e, "hipGetDeviceCount");
    }
    *count = n;
    if (n == 0) {
        set_error("no HIP device");
        return PSW_ERR_DEVICE;
    }
    return PSW_OK;
}

int psw_set_device(int device) {
    PSW_TRY_HIP(hipSetDevice(device), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_device_name(char* buf, size_t n) {
    PSW_CHECK_ARG(buf && n > 0, "buffer is NULL");
    int dev = 0;
    PSW_TRY_HIP(hipGetDevice(&dev), PSW_ERR_DEVICE);
    hipDeviceProp_t prop;
    PSW_TRY_HIP(hipGetDeviceProperties(&prop, dev), PSW_ERR_DEVICE);
    snprintf(buf, n, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return PSW_OK;
}

int psw_malloc(void** d_ptr, size_t bytes) {
    PSW_CHECK_ARG(d_ptr, "d_ptr is NULL");
    *d_ptr = nullptr;
    if (bytes == 0) return PSW_OK;
    PSW_TRY_HIP(hipMalloc(d_ptr, bytes), PSW_ERR_ALLOC);
    return PSW_OK;
}

int psw_free(void* d_ptr) {
    if (!d_ptr) return PSW_OK;
    PSW_TRY_HIP(hipFree(d_ptr), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_host_alloc(void** h_ptr, size_t bytes) {
    PSW_CHECK_ARG(h_ptr, "h_ptr is NULL");
    *h_ptr = nullptr;
    if (bytes == 0) return PSW_OK;
    PSW_TRY_HIP(hipHostMalloc(h_ptr, bytes, hipHostMallocDefault), PSW_ERR_ALLOC);
    return PSW_OK;
}

int psw_host_free(void* h_ptr) {
    if (!h_ptr) return PSW_OK;
    PSW_TRY_HIP(hipHostFree(h_ptr), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_memcpy_htod(void* d_dst, const void* src, size_t bytes, psw_stream_t s) {
    if (bytes == 0) return PSW_OK;
    PSW_CHECK_ARG(d_dst && src, "NULL pointer");
    PSW_TRY_HIP(hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, S(s)), PSW_ERR_COPY);
    PSW_TRY_HIP(hipStreamSynchronize(S(s)), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_memcpy_dtoh(void* dst, const void* d_src, size_t bytes, psw_stream_t s) {
    if (bytes == 0) return PSW_OK;
    PSW_CHECK_ARG(dst && d_src, "NULL pointer");
    PSW_TRY_HIP(hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, S(s)), PSW_ERR_COPY);
    PSW_TRY_HIP(hipStreamSynchronize(S(s)), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_memcpy_dtod(void* d_dst, const void* d_src, size_t bytes, psw_stream_t s) {
    if (bytes == 0) return PSW_OK;
    PSW_CHECK_ARG(d_dst && d_src, "NULL pointer");
    PSW_TRY_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, S(s)), PSW_ERR_COPY);
    return PSW_OK;
}

int psw_fill_f32(float* d_ptr, float value, size_t count, psw_stream_t s) {
    if (count == 0) return PSW_OK;
    PSW_CHECK_ARG(d_ptr, "d_ptr is NULL");
    fill_f32_kernel<<<fill_blocks(count), 256, 0, S(s)>>>(d_ptr, value, count);
    return launch_status("fill_f32_kernel");
}

int psw_fill_u32(uint32_t* d_ptr, uint32_t value, size_t count, psw_stream_t s) {
    if (count == 0) return PSW_OK;
    PSW_CHECK_ARG(d_ptr, "d_ptr is NULL");
    fill_u32_kernel<<<fill_blocks(count), 256, 0, S(s)>>>(d_ptr, value, count);
    return launch_status("fill_u32_kernel");
}

int psw_copy_f32(float* d_dst, const float* d_src, size_t count, psw_stream_t s) {
    if (count == 0) return PSW_OK;
    PSW_CHECK_ARG(d_dst && d_src, "NULL pointer");
    PSW_CHECK_ARG(count % 4 == 0 && ((uintptr_t)d_dst % 16) == 0 && ((uintptr_t)d_src % 16) == 0,
                  "count must be a multiple of 4 and pointers 16-byte aligned");
    copy_f32_kernel<<<fill_blocks(count / 4), 256, 0, S(s)>>>(reinterpret_cast<float4*>(d_dst),
                                                             reinterpret_cast<const float4*>(d_src), count / 4);
    return launch_status("copy_f32_kernel");
}

int psw_stream_create(psw_stream_t* s) {
    PSW_CHECK_ARG(s, "s is NULL");
    hipStream_t st;
    PSW_TRY_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), PSW_ERR_DEVICE);
    *s = st;
    return PSW_OK;
}

int psw_stream_destroy(psw_stream_t s) {
    if (!s) return PSW_OK;
    PSW_TRY_HIP(hipStreamDestroy(S(s)), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_stream_synchronize(psw_stream_t s) {
    PSW_TRY_HIP(hipStreamSynchronize(S(s)), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_device_synchronize(void) {
    PSW_TRY_HIP(hipDeviceSynchronize(), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_event_create(psw_event_t* e) {
    PSW_CHECK_ARG(e, "e is NULL");
    hipEvent_t ev;
    PSW_TRY_HIP(hipEventCreate(&ev), PSW_ERR_DEVICE);
    *e = ev;
    return PSW_OK;
}

int psw_event_destroy(psw_event_t e) {
    if (!e) return PSW_OK;
    PSW_TRY_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(e)), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_event_record(psw_event_t e, psw_stream_t s) {
    PSW_CHECK_ARG(e, "e is NULL");
    PSW_TRY_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(e), S(s)), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_event_synchronize(psw_event_t e) {
    PSW_CHECK_ARG(e, "e is NULL");
    PSW_TRY_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(e)), PSW_ERR_DEVICE);
    return PSW_OK;
}

int psw_event_elapsed_ms(float* ms, psw_event_t start, psw_event_t stop) {
    PSW_CHECK_ARG(ms && start && stop, "NULL argument");
    PSW_TRY_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start), reinterpret_cast<hipEvent_t>(stop)),
                PSW_ERR_DEVICE);
    return PSW_OK;
}

}  // extern "C"
