// Runtime plumbing of librtstereo: error reporting, device memory, streams,
// events, layout conversion between the reference's pixel-major volumes and
// the disparity-major device layout, and small utility kernels.
#include "rts_internal.h"

#include <stdarg.h>
#include <stdio.h>

namespace rts {

static thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

int check_dims(int W, int H, int D) {
    RTS_CHECK(W >= 1 && H >= 1, "image size %dx%d must be positive", W, H);
    RTS_CHECK(D >= 1 && D <= 256, "disparity range %d must be in [1, 256]", D);
    RTS_CHECK(D <= W, "disparity range %d exceeds the image width %d", D, W);
    RTS_CHECK((size_t)W * H <= (size_t)1 << 30, "image of %dx%d pixels is too large", W, H);
    return RTS_OK;
}

// Pixel-major [P][D] -> disparity-major [D][P] through a 64x64 LDS tile.
// Block (64, 4): tile columns are 64 consecutive elements of the source rows.
__global__ void __launch_bounds__(256) k_transpose(const float *__restrict__ src, float *__restrict__ dst,
                                                   long rows, long cols) {
    // src is [rows][cols]; dst is [cols][rows].
    __shared__ float tile[64][65];
    const long r0 = (long)blockIdx.x * 64, c0 = (long)blockIdx.y * 64;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int i = ty; i < 64; i += 4) {
        const long r = r0 + i, c = c0 + tx;
        if (r < rows && c < cols) tile[i][tx] = src[r * cols + c];
    }
    __syncthreads();
    for (int i = ty; i < 64; i += 4) {
        const long c = c0 + i, r = r0 + tx;
        if (r < rows && c < cols) dst[c * rows + r] = tile[tx][i];
    }
}

__global__ void k_fill_u8(uint8_t *dst, uint8_t v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

__global__ void __launch_bounds__(256) k_copy16(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

__global__ void k_rgb_to_gray(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ gray, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = rgb[3 * i], g = rgb[3 * i + 1], r = rgb[3 * i + 2];
    gray[i] = (uint8_t)((29u * b + 150u * g + 77u * r + 128u) >> 8);
}

}  // namespace rts

using namespace rts;

extern "C" {

const char *rts_last_error(void) { return rts::g_err; }
const char *rts_version(void) { return "rtstereo-mi355x 0.1 (gfx950)"; }

rts_status rts_device_count(int *n) {
    RTS_HIP(hipGetDeviceCount(n));
    return RTS_OK;
}
rts_status rts_set_device(int device) {
    RTS_HIP(hipSetDevice(device));
    return RTS_OK;
}
rts_status rts_malloc(void **ptr, size_t bytes) {
    RTS_CHECK(ptr != nullptr, "rts_malloc: null out pointer");
    *ptr = nullptr;
    if (bytes == 0) return RTS_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *ptr = nullptr;
        return fail(RTS_ERR_NOMEM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return RTS_OK;
}
rts_status rts_free(void *ptr) {
    if (ptr) RTS_HIP(hipFree(ptr));
    return RTS_OK;
}
rts_status rts_host_alloc(void **ptr, size_t bytes) {
    RTS_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return RTS_OK;
}
rts_status rts_host_free(void *ptr) {
    if (ptr) RTS_HIP(hipHostFree(ptr));
    return RTS_OK;
}
rts_status rts_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes) RTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(stream)));
    return RTS_OK;
}
rts_status rts_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes) RTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(stream)));
    return RTS_OK;
}
rts_status rts_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes) RTS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(stream)));
    return RTS_OK;
}
rts_status rts_fill_u8(void *dst, uint8_t value, size_t bytes, void *stream) {
    if (!bytes) return RTS_OK;
    const int blocks = (int)std::min<size_t>((bytes + 255) / 256, 4096);
    k_fill_u8<<<blocks, 256, 0, S(stream)>>>((uint8_t *)dst, value, bytes);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_stream_create(void **stream) {
    hipStream_t s;
    RTS_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (void *)s;
    return RTS_OK;
}
rts_status rts_stream_destroy(void *stream) {
    if (stream) RTS_HIP(hipStreamDestroy(S(stream)));
    return RTS_OK;
}
rts_status rts_stream_sync(void *stream) {
    RTS_HIP(hipStreamSynchronize(S(stream)));
    return RTS_OK;
}
rts_status rts_device_sync(void) {
    RTS_HIP(hipDeviceSynchronize());
    return RTS_OK;
}
rts_status rts_event_create(void **ev) {
    hipEvent_t e;
    RTS_HIP(hipEventCreate(&e));
    *ev = (void *)e;
    return RTS_OK;
}
rts_status rts_event_destroy(void *ev) {
    if (ev) RTS_HIP(hipEventDestroy((hipEvent_t)ev));
    return RTS_OK;
}
rts_status rts_event_record(void *ev, void *stream) {
    RTS_HIP(hipEventRecord((hipEvent_t)ev, S(stream)));
    return RTS_OK;
}
rts_status rts_event_sync(void *ev) {
    RTS_HIP(hipEventSynchronize((hipEvent_t)ev));
    return RTS_OK;
}
rts_status rts_event_elapsed_ms(void *start, void *stop, float *ms) {
    RTS_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return RTS_OK;
}
rts_status rts_mem_info(size_t *free_bytes, size_t *total_bytes) {
    RTS_HIP(hipMemGetInfo(free_bytes, total_bytes));
    return RTS_OK;
}

rts_status rts_vol_to_dm(const float *pm, float *dm, int W, int H, int D, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    const long P = (long)W * H;
    dim3 grid(div_up(P, 64), div_up(D, 64)), block(64, 4);
    k_transpose<<<grid, block, 0, S(stream)>>>(pm, dm, P, D);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_vol_to_pm(const float *dm, float *pm, int W, int H, int D, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    const long P = (long)W * H;
    RTS_CHECK(div_up(P, 64) <= 65535, "volume of %ld pixels is too large to convert", P);
    dim3 grid(div_up(D, 64), div_up(P, 64)), block(64, 4);
    k_transpose<<<grid, block, 0, S(stream)>>>(dm, pm, D, P);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_copy_kernel(const void *src, void *dst, size_t bytes, void *stream) {
    RTS_CHECK(bytes % 16 == 0, "copy size %zu is not a multiple of 16", bytes);
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 64);
    k_copy16<<<blocks, 256, 0, S(stream)>>>((const uint4 *)src, (uint4 *)dst, n);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_rgb_to_gray(const uint8_t *rgb, uint8_t *gray, int W, int H, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size");
    const int n = W * H;
    k_rgb_to_gray<<<div_up(n, 256), 256, 0, S(stream)>>>(rgb, gray, n);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
