// Adaptive-weight aggregation with the support weights built on the fly
// from the two Lab images (seq/AdaptiveWeightAggregation.cpp:11-111 without
// the per-pixel weight tensors, which would be 9 GB per view at 1080p).
//
// For pixel p, disparity d and target centre q = (x-d wrapped, y):
//   w_ref(p,k) = prox(k) * exp(-|Lab_L(p) - Lab_L(p+k)| / gc)   (0 outside)
//   w_tar(q,k) = prox(k) * exp(-|Lab_R(q) - Lab_R(q+k)| / gc)   (0 outside)
//   out(p,d)   = sum_k w_ref*w_tar*C(p+k wrapped, d) / sum_k w_ref*w_tar
// in the reference's row-major window order.
//
// Block: 64 consecutive pixels of one row x AW_DC disparities.  For each
// window row dy the block stages in LDS the ref weights of its 64 pixels,
// the tar weights of the 64+AW_DC-1 target centres it touches and the cost
// rows of its disparities; each thread then accumulates one pixel x AW_DPT
// disparities across the window row.  FP32 keeps the reference's
// arithmetic (product, add, product, add); FP16 stores the weights in half,
// forms w_ref*w_tar and C*w in half and accumulates in float.
#include "rts_internal.h"

#include <hip/hip_fp16.h>

#pragma clang fp contract(off)

namespace rts {

constexpr int AW_TX = 64;
constexpr int AW_DPT = 4;
constexpr int AW_DC = 16;               // 4 thread rows x AW_DPT
constexpr int AW_RMAX = 16;             // window radius limit of this kernel
constexpr int AW_E = 2 * AW_RMAX + 1;   // 33
constexpr int AW_QW = AW_TX + AW_DC - 1;

template <typename T>
struct AwLds {
    T el[AW_E][AW_TX];
    T er[AW_E][AW_QW + 1];
    float c[AW_DC][AW_TX + 2 * AW_RMAX + 1];
};

__device__ __forceinline__ float lab_dist3(const uint8_t *a, const uint8_t *b) {
    const float e0 = (float)a[0] - (float)b[0], e1 = (float)a[1] - (float)b[1], e2 = (float)a[2] - (float)b[2];
    return sqrt_rn(e0 * e0 + e1 * e1 + e2 * e2);
}

template <bool HALF>
__global__ void __launch_bounds__(256) k_aw(const uint8_t *__restrict__ labL, const uint8_t *__restrict__ labR,
                                            const float *__restrict__ vol, float *__restrict__ out,
                                            const float *__restrict__ prox, int W, int H, int D, int r, float gc) {
    using T = typename std::conditional<HALF, __half, float>::type;
    __shared__ AwLds<T> sm;
    const int x0 = blockIdx.x * AW_TX, y = blockIdx.y, d0 = blockIdx.z * AW_DC;
    const int lx = threadIdx.x & 63, dg = threadIdx.x >> 6;
    const int x = x0 + lx;
    const int e = 2 * r + 1;
    const size_t pl = plane(W, H);
    float sw[AW_DPT], sc[AW_DPT];
#pragma unroll
    for (int j = 0; j < AW_DPT; j++) sw[j] = sc[j] = 0.0f;
    const int qbase = x0 - (d0 + AW_DC - 1);   // unwrapped x of er column 0
    for (int dy = -r; dy <= r; dy++) {
        const int yy = y + dy;   // weights: no wrap, zero outside
        const int yc = wrap1(yy, H);   // costs: wrapped
        __syncthreads();
        // Ref weights el[dx][i] for pixel x0+i.
        for (int t = threadIdx.x; t < e * AW_TX; t += 256) {
            const int dxi = t / AW_TX, i = t - dxi * AW_TX;
            const int px = x0 + i, xx = px + dxi - r;
            float v = 0.0f;
            if (px < W && yy >= 0 && yy < H && xx >= 0 && xx < W) {
                const float dist = lab_dist3(labL + 3 * ((size_t)y * W + px), labL + 3 * ((size_t)yy * W + xx));
                v = prox[(dy + r) * e + dxi] * expf(__fdiv_rn(-dist, gc));
            }
            sm.el[dxi][i] = (T)v;
        }
        // Tar weights er[dx][j] for target centre (qbase+j) wrapped.
        for (int t = threadIdx.x; t < e * AW_QW; t += 256) {
            const int dxi = t / AW_QW, j = t - dxi * AW_QW;
            int qx = qbase + j;
            qx = qx < 0 ? qx + W : (qx >= W ? qx - W : qx);
            float v = 0.0f;
            if (qx >= 0 && qx < W) {
                const int xx = qx + dxi - r;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const float dist = lab_dist3(labR + 3 * ((size_t)y * W + qx), labR + 3 * ((size_t)yy * W + xx));
                    v = prox[(dy + r) * e + dxi] * expf(__fdiv_rn(-dist, gc));
                }
            }
            sm.er[dxi][j] = (T)v;
        }
        // Cost rows c[dd][i] = C(x0 - r + i wrapped, yc, d0 + dd).
        for (int t = threadIdx.x; t < AW_DC * (AW_TX + 2 * r); t += 256) {
            const int dd = t / (AW_TX + 2 * r), i = t - dd * (AW_TX + 2 * r);
            const int d = d0 + dd;
            float v = 0.0f;
            if (d < D) v = vol[(size_t)d * pl + (size_t)yc * W + wrapn(x0 - r + i, W)];
            sm.c[dd][i] = v;
        }
        __syncthreads();
        if (x < W) {
            for (int dxi = 0; dxi < e; dxi++) {
                const T wl = sm.el[dxi][lx];
#pragma unroll
                for (int j = 0; j < AW_DPT; j++) {
                    const int dd = dg * AW_DPT + j;
                    // target centre x - d -> er column lx - dd + AW_DC - 1
                    const T wt = sm.er[dxi][lx - dd + AW_DC - 1];
                    const float c = sm.c[dd][lx + dxi];
                    if (HALF) {
                        const __half w = __hmul(wl, wt);
                        const __half cw = __hmul(__float2half(c), w);
                        sw[j] += __half2float(w);
                        sc[j] += __half2float(cw);
                    } else {
                        const float w = (float)wl * (float)wt;
                        sw[j] += w;
                        const float cw = c * w;
                        sc[j] += cw;
                    }
                }
            }
        }
    }
    if (x >= W) return;
#pragma unroll
    for (int j = 0; j < AW_DPT; j++) {
        const int d = d0 + dg * AW_DPT + j;
        if (d < D) out[(size_t)d * pl + (size_t)y * W + x] = __fdiv_rn(sc[j], sw[j]);
    }
}

}  // namespace rts

using namespace rts;

extern "C" rts_status rts_aw(const uint8_t *lab_ref, const uint8_t *lab_tar, const float *vol, float *out, int W,
                             int H, int D, int r, float gamma_c, float gamma_p, int precision, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r <= AW_RMAX && r < W && r < H, "AW radius %d must be in [0, %d] and fit the image", r,
              AW_RMAX);
    RTS_CHECK(vol != out, "AW aggregation is out of place");
    RTS_CHECK(precision == RTS_AW_FP32 || precision == RTS_AW_FP16, "bad AW precision %d", precision);
    RTS_CHECK(H <= 65535, "height %d too large", H);
    hipStream_t s = S(stream);
    float *prox = nullptr;
    RTS_HIP(hipMallocAsync((void **)&prox, sizeof(float) * (2 * r + 1) * (2 * r + 1), s));
    int st = rts_prox_weight(prox, gamma_p, r, s);
    if (st == RTS_OK) {
        dim3 grid(div_up(W, AW_TX), H, div_up(D, AW_DC));
        if (precision == RTS_AW_FP16)
            k_aw<true><<<grid, 256, 0, s>>>(lab_ref, lab_tar, vol, out, prox, W, H, D, r, gamma_c);
        else
            k_aw<false><<<grid, 256, 0, s>>>(lab_ref, lab_tar, vol, out, prox, W, H, D, r, gamma_c);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = fail(RTS_ERR_HIP, "AW launch: %s", hipGetErrorString(e));
    }
    (void)hipFreeAsync(prox, s);
    return st;
}
