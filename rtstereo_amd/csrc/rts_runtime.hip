// Runtime plumbing of librtstereo: error reporting, device memory, streams,
// events, layout conversion between the reference's pixel-major volumes and
// the disparity-major device layout, and small utility kernels.
#include "rts_internal.h"

#include <math.h>

#include <mutex>

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

// FP32 -> FP16 volume, each value rounded to nearest: a thread converts
// groups of 8 floats into one 16-byte word.
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_to_half(const float *__restrict__ src, u32x4 *__restrict__ dst, size_t n8) {
    const f32x4_t *s4 = reinterpret_cast<const f32x4_t *>(src);
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n8; i += (size_t)gridDim.x * 256) {
        const f32x4_t a = __builtin_nontemporal_load(s4 + 2 * i), b = __builtin_nontemporal_load(s4 + 2 * i + 1);
        u32x4 o;
        o.x = __builtin_bit_cast(unsigned int, h2_t{(_Float16)a.x, (_Float16)a.y});
        o.y = __builtin_bit_cast(unsigned int, h2_t{(_Float16)a.z, (_Float16)a.w});
        o.z = __builtin_bit_cast(unsigned int, h2_t{(_Float16)b.x, (_Float16)b.y});
        o.w = __builtin_bit_cast(unsigned int, h2_t{(_Float16)b.z, (_Float16)b.w});
        dst[i] = o;
    }
}

// Copy probe: a block moves rounds of 256 x CP_U 16-byte words, each thread
// issuing all CP_U loads of a round before its first store (CP_U loads in
// flight per lane); loads and stores are non-temporal, as a stream's are.
constexpr int CP_U = 8;
__global__ void __launch_bounds__(256) k_copy16(const u32x4 *__restrict__ src, u32x4 *__restrict__ dst, size_t n) {
    constexpr size_t per = 256 * CP_U;
    for (size_t b = (size_t)blockIdx.x * per; b < n; b += (size_t)gridDim.x * per) {
        u32x4 v[CP_U];
#pragma unroll
        for (int k = 0; k < CP_U; k++) {
            const size_t i = b + (size_t)k * 256 + threadIdx.x;
            if (i < n) v[k] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int k = 0; k < CP_U; k++) {
            const size_t i = b + (size_t)k * 256 + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(v[k], dst + i);
        }
    }
}

// Read and write probes in the copy probe's shape (rounds of 256 x CP_U
// 16-byte words per block, all CP_U accesses of a round in flight per lane):
// the read probe folds what it reads into one word per block, so that the
// loads are kept; the write probe stores a pattern.
__global__ void __launch_bounds__(256) k_read16(const u32x4 *__restrict__ src, size_t n, uint32_t *__restrict__ sink) {
    constexpr size_t per = 256 * CP_U;
    uint32_t acc = 0;
    for (size_t b = (size_t)blockIdx.x * per; b < n; b += (size_t)gridDim.x * per) {
        u32x4 v[CP_U];
#pragma unroll
        for (int k = 0; k < CP_U; k++) {
            const size_t i = b + (size_t)k * 256 + threadIdx.x;
            v[k] = i < n ? __builtin_nontemporal_load(src + i) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int k = 0; k < CP_U; k++) acc ^= v[k].x ^ v[k].y ^ v[k].z ^ v[k].w;
    }
    // (a store that depends on every load, so that none is dropped; the
    // sink holds one word per block)
    if (acc == 0x9E3779B9u) sink[blockIdx.x] = acc;
}

__global__ void __launch_bounds__(256) k_write16(u32x4 *__restrict__ dst, size_t n) {
    constexpr size_t per = 256 * CP_U;
    for (size_t b = (size_t)blockIdx.x * per; b < n; b += (size_t)gridDim.x * per) {
#pragma unroll
        for (int k = 0; k < CP_U; k++) {
            const size_t i = b + (size_t)k * 256 + threadIdx.x;
            if (i < n) __builtin_nontemporal_store(u32x4{(uint32_t)i, 1u, 2u, 3u}, dst + i);
        }
    }
}

__global__ void k_rgb_to_gray(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ gray, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gray[i] = gray_bgr(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
}

// 8-bit CIELab (D65) from interleaved B,G,R bytes, in FP32: sRGB
// linearised, XYZ, then L*255/100, a+128, b+128, rounded half to even and
// clipped.  (The bytes equal an FP64 evaluation except where a value lies
// within FP32 rounding of a half-integer: about 1.5e-5 of them, by one level.)
__device__ __forceinline__ float srgb_lin(uint8_t v) {
    const float c = (float)v / 255.0f;
    return c <= 0.04045f ? c / 12.92f : powf((c + 0.055f) / 1.055f, 2.4f);
}

__device__ __forceinline__ float lab_f(float t) {
    const float eps = 216.0f / 24389.0f, kap = 24389.0f / 27.0f;
    return t > eps ? cbrtf(t) : (kap * t + 16.0f) / 116.0f;
}

__global__ void k_rgb_to_lab(const uint8_t *__restrict__ rgb, uint8_t *__restrict__ lab, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float b = srgb_lin(rgb[3 * i]), g = srgb_lin(rgb[3 * i + 1]), r = srgb_lin(rgb[3 * i + 2]);
    const float X = (0.4124564f * r + 0.3575761f * g + 0.1804375f * b) / 0.950456f;
    const float Y = 0.2126729f * r + 0.7151522f * g + 0.0721750f * b;
    const float Z = (0.0193339f * r + 0.1191920f * g + 0.9503041f * b) / 1.088754f;
    const float fx = lab_f(X), fy = lab_f(Y), fz = lab_f(Z);
    const float v[3] = {(116.0f * fy - 16.0f) * 255.0f / 100.0f, 500.0f * (fx - fy) + 128.0f,
                        200.0f * (fy - fz) + 128.0f};
#pragma unroll
    for (int c = 0; c < 3; c++) lab[3 * i + c] = (uint8_t)fminf(fmaxf(__builtin_rintf(v[c]), 0.0f), 255.0f);
}

// 3x3 median of each colour channel with replicated borders (the OpenCL
// driver's pre-filter, opencl/Main.cpp:342-345, there from OpenCV's
// medianBlur).  A 19-exchange sorting network on the nine values.
__device__ __forceinline__ void cswap(int &a, int &b) {
    const int lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

__global__ void k_median3(const uint8_t *__restrict__ in, uint8_t *__restrict__ out, int W, int H) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    for (int c = 0; c < 3; c++) {
        int v[9], k = 0;
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int yy = min(max(y + dy, 0), H - 1), xx = min(max(x + dx, 0), W - 1);
                v[k++] = in[3 * ((size_t)yy * W + xx) + c];
            }
        cswap(v[1], v[2]); cswap(v[4], v[5]); cswap(v[7], v[8]);
        cswap(v[0], v[1]); cswap(v[3], v[4]); cswap(v[6], v[7]);
        cswap(v[1], v[2]); cswap(v[4], v[5]); cswap(v[7], v[8]);
        cswap(v[0], v[3]); cswap(v[5], v[8]); cswap(v[4], v[7]);
        cswap(v[3], v[6]); cswap(v[1], v[4]); cswap(v[2], v[5]);
        cswap(v[4], v[7]); cswap(v[4], v[2]); cswap(v[6], v[4]);
        cswap(v[4], v[2]);
        out[3 * ((size_t)y * W + x) + c] = (uint8_t)v[4];
    }
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

rts_status rts_volume_to_half(const float *src, void *dst, size_t n, void *stream) {
    RTS_CHECK(n % 8 == 0, "volume of %zu floats is not a multiple of 8", n);
    RTS_CHECK(((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0,
              "volumes must be 16-byte aligned");
    const size_t n8 = n / 8;
    const int blocks = (int)std::min<size_t>((n8 + 255) / 256, 256 * 16);
    k_to_half<<<blocks, 256, 0, S(stream)>>>(src, static_cast<u32x4 *>(dst), n8);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_copy_kernel(const void *src, void *dst, size_t bytes, void *stream) {
    RTS_CHECK(bytes % 16 == 0, "copy size %zu is not a multiple of 16", bytes);
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 256 * CP_U - 1) / (256 * CP_U), 256 * 16);
    k_copy16<<<blocks, 256, 0, S(stream)>>>((const u32x4 *)src, (u32x4 *)dst, n);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_read_kernel(const void *src, size_t bytes, void *sink, void *stream) {
    RTS_CHECK(bytes % 16 == 0, "read size %zu is not a multiple of 16", bytes);
    RTS_CHECK(sink != nullptr, "the read probe needs a sink of %d words", 256 * 16);
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 256 * CP_U - 1) / (256 * CP_U), 256 * 16);
    k_read16<<<blocks, 256, 0, S(stream)>>>((const u32x4 *)src, n, (uint32_t *)sink);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_write_kernel(void *dst, size_t bytes, void *stream) {
    RTS_CHECK(bytes % 16 == 0, "write size %zu is not a multiple of 16", bytes);
    const size_t n = bytes / 16;
    const int blocks = (int)std::min<size_t>((n + 256 * CP_U - 1) / (256 * CP_U), 256 * 16);
    k_write16<<<blocks, 256, 0, S(stream)>>>((u32x4 *)dst, n);
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

rts_status rts_rgb_to_lab(const uint8_t *rgb, uint8_t *lab, int W, int H, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && (long)W * H < (1L << 31), "bad image size %dx%d", W, H);
    const int n = W * H;
    k_rgb_to_lab<<<div_up(n, 256), 256, 0, S(stream)>>>(rgb, lab, n);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_median3(const uint8_t *rgb, uint8_t *out, int W, int H, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && div_up(H, 4) <= 65535, "bad image size %dx%d", W, H);
    RTS_CHECK(rgb != out, "rts_median3 is out of place");
    dim3 grid(div_up(W, 64), div_up(H, 4)), block(64, 4);
    k_median3<<<grid, block, 0, S(stream)>>>(rgb, out, W, H);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
