// Internal helpers shared by the HIP translation units of librtstereo.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "rtstereo.h"
#include "sqrt_rn.h"

namespace rts {

// Records a message for rts_last_error() (thread-local) and returns `code`.
int fail(int code, const char *fmt, ...);

// Maps a HIP error to RTS_ERR_HIP with the call site recorded.
#define RTS_HIP(call)                                                                   \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::rts::fail(RTS_ERR_HIP, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                               __FILE__, __LINE__);                                     \
    } while (0)

#define RTS_CHECK(cond, ...)                                                            \
    do {                                                                                \
        if (!(cond)) return ::rts::fail(RTS_ERR_INVALID, __VA_ARGS__);                  \
    } while (0)

// After a kernel launch: report a launch failure as RTS_ERR_HIP.
#define RTS_LAUNCHED()                                                                  \
    do {                                                                                \
        hipError_t e_ = hipGetLastError();                                              \
        if (e_ != hipSuccess)                                                           \
            return ::rts::fail(RTS_ERR_HIP, "launch: %s (%s:%d)", hipGetErrorString(e_),  \
                               __FILE__, __LINE__);                                     \
    } while (0)

inline hipStream_t S(void *stream) { return reinterpret_cast<hipStream_t>(stream); }

inline int div_up(long a, long b) { return (int)((a + b - 1) / b); }

// One-step toroidal wrap of an index that is at most one period out of range
// (the reference's border rule for every window and for x-d).
__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Full modulo for tile-staging indices that may lie several periods away on
// narrow images (only in-period entries are ever consumed).
__device__ __forceinline__ int wrapn(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// Correctly rounded float square root.  The hardware v_sqrt_f32 (which
// __fsqrt_rn lowers to here) is within one ulp but not correctly rounded;
// rts_sqrt_fix moves its result to RN(sqrt(x)) with exact integer midpoint
// tests (sqrt_rn.h; no double precision).
__device__ __forceinline__ float hw_sqrt(float v) { return __builtin_amdgcn_sqrtf(v); }
__device__ __forceinline__ float sqrt_rn(float x) { return rts_sqrt_rn_with(x, hw_sqrt); }

// Store of a volume element that the next kernel reads once: non-temporal,
// so streamed volumes do not displace reusable lines from the caches.
__device__ __forceinline__ void stream_store(float *p, float v) { __builtin_nontemporal_store(v, p); }

// a / c for a count c (a window area or support size) in [1, 2^19], rounded as IEEE division rounds
// it.  With y = rcp(c) (relative error e, |e| <= 2^-22), q0 = a*y is within a
// few ulp of a/c; the residual r0 = a - c*q0 is then exact (a multiple of
// ulp(q0) no larger than 2^24 of them, since c is an integer), and
// q0 + r0*y = a/c + (a/c - q0)*e lies within 2^-19 ulp of a/c.  A quotient a/c
// of a float by an integer c is never a rounding midpoint in the normal range
// and lies at least ulp/(2c) from one, so the fused q1 = RN(q0 + r0*y) is
// RN(a/c).  Results that come out denormal, infinite or NaN (where midpoints
// do occur, or the residual is not finite) are recomputed by the division,
// and so are zero results: for a = -0 the fused steps give +0 where the
// division gives -0.
__device__ __forceinline__ float div_count(float a, uint32_t c) {
    const float cf = (float)c;
    const float y = __builtin_amdgcn_rcpf(cf);
    const float q0 = a * y;
    const float r0 = __builtin_fmaf(-q0, cf, a);
    const float q1 = __builtin_fmaf(r0, y, q0);
    // classes: signalling/quiet NaN, -inf, -denormal, -0, +0, +denormal, +inf
    if (__builtin_expect(__builtin_amdgcn_classf(q1, 0x2F7), 0)) return __fdiv_rn(a, cf);
    return q1;
}

// Word 3 of a raw buffer resource on CDNA (gfx9 family): 32-bit data format,
// no swizzle, no stride, so an access is base + soffset + voffset.
constexpr int BUF_DWORD3 = 0x00020000;

// XCD-aware block order.  MI355X hands consecutive workgroups of a launch
// to its 8 XCDs round-robin, and each XCD has its own L2.  Given the linear
// workgroup id, return a work index such that XCD k processes one
// contiguous range of the work sequence: neighbouring work items then share
// an L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_order(int lin, int n) {
    const int per = n >> 3, rem = n & 7;
    const int xcd = lin & 7, j = lin >> 3;
    return xcd < rem ? xcd * (per + 1) + j : rem * (per + 1) + (xcd - rem) * per + j;
}

// Plane stride of a disparity-major volume [D][H][W].
__host__ __device__ __forceinline__ size_t plane(int W, int H) { return (size_t)W * (size_t)H; }

// Validates W, H, D against the limits every entry point shares.
int check_dims(int W, int H, int D);

// Winner-takes-all (rts_wta's work) with the caller's per-row scratch for the
// literal carry (`carry`: H ints; unused, and may be NULL, without
// RTS_WTA_LITERAL).  Rows at most WTA_MAX_ROWS.
constexpr int WTA_MAX_ROWS = 65535;
int wta_launch(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags, int tol,
               int *carry, hipStream_t s);

// rts_pipeline_create's checks of a configuration, without allocating.
int pipeline_check(const rts_pipeline_desc &d);

// One run's launch sequence of a pipeline on stream s, without graph
// replay (colour images in; gray and Lab derived on the device): what
// rts_pipeline_run launches for those buffers, for callers that capture the
// launches of many pipelines into one graph (the batch engine).
int pipeline_enqueue(rts_pipeline *p, const uint8_t *left_rgb, const uint8_t *right_rgb, uint8_t *dl, uint8_t *dr,
                     uint8_t *mask, hipStream_t s);

// Devices a process can address through per-device tables.
constexpr int RTS_MAX_DEVICES = 64;

// Cross arms (arm limit at most 17) from a packed 0x00BBGGRR colour image.
int find_cross_packed(const uint32_t *pcol, uint32_t *arms, int W, int H, int L, int tao, hipStream_t s);

// Colour AD helper: |a-b| summed over the three bytes of packed 0x00BBGGRR
// pixels (v_sad_u8 does four byte |differences| plus an accumulator).
__device__ __forceinline__ uint32_t sad3(uint32_t a, uint32_t b) { return __builtin_amdgcn_sad_u8(a, b, 0u); }

__device__ __forceinline__ uint32_t pack_rgb(const uint8_t *p) {
    return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16);
}

// Maximum per-channel |difference| of two packed pixels.
__device__ __forceinline__ int max_chan_delta(uint32_t a, uint32_t b) {
    int m = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        int t = (int)((a >> (8 * c)) & 255u) - (int)((b >> (8 * c)) & 255u);
        t = t < 0 ? -t : t;
        m = t > m ? t : m;
    }
    return m;
}

// Gray level of a B,G,R pixel: fixed-point weights 29/150/77 over 256,
// rounded (OpenCV's BGR2GRAY integer form).
__device__ __forceinline__ uint8_t gray_bgr(uint32_t b, uint32_t g, uint32_t r) {
    return (uint8_t)((29u * b + 150u * g + 77u * r + 128u) >> 8);
}

// Packed arm layout: byte 0 left, 1 right, 2 top, 3 bottom.
__device__ __forceinline__ int arm_l(uint32_t a) { return (int)(a & 255u); }
__device__ __forceinline__ int arm_r(uint32_t a) { return (int)((a >> 8) & 255u); }
__device__ __forceinline__ int arm_t(uint32_t a) { return (int)((a >> 16) & 255u); }
__device__ __forceinline__ int arm_b(uint32_t a) { return (int)(a >> 24); }

}  // namespace rts
