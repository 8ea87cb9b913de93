// Disparity selection: winner-takes-all with the right view read off the
// volume's diagonal and the left-right check fused in, the stand-alone
// cross-check, and the cross-support histogram refinement.
#include "rts_internal.h"

namespace rts {

// One block per image row; thread t owns pixels t, t+256, ...  For every d
// it reads vol[d][y][x] for the left view and vol[d][y][(x+d) mod W] for the
// right view (right pixel x matched with left pixel x+d); both reads are
// contiguous across the wavefront, and the second mostly hits the cache
// lines the first brought in.  Strict comparisons from d=0 up keep the
// lowest d on ties (seq/WTA.cpp:16-21).  WTA_DU disparities are loaded
// before they are compared, so their loads are in flight together.  After
// the sweep the row's D_L and D_R go to LDS and the left-right check
// (seq/CrossCheck.cpp:15-21) runs there.
constexpr int WTA_THREADS = 256;
constexpr int WTA_DU = 4;

template <int PPT, bool MAXIMIZE>
__global__ void __launch_bounds__(WTA_THREADS) k_wta(const float *__restrict__ vol, uint8_t *__restrict__ dl,
                                                     uint8_t *__restrict__ dr, uint8_t *__restrict__ mask,
                                                     uint8_t *__restrict__ beat0, int W, int H, int D, int tol) {
    extern __shared__ uint8_t maps[];   // 2*W bytes: D_L then D_R of the row
    uint8_t *mapL = maps, *mapR = maps + W;
    const int y = blockIdx.x;
    const size_t pl = plane(W, H), rowoff = (size_t)y * W;
    const bool want_r = dr != nullptr;
    const float *row = vol + rowoff;
    float bestL[PPT], bestR[PPT];
    int dL[PPT], dR[PPT], xs[PPT];
#pragma unroll
    for (int i = 0; i < PPT; i++) {
        xs[i] = min((int)threadIdx.x + i * WTA_THREADS, W - 1);
        bestL[i] = bestR[i] = 0.0f;
        dL[i] = dR[i] = 0;
    }
    for (int d0 = 0; d0 < D; d0 += WTA_DU) {
        float vL[WTA_DU][PPT], vR[WTA_DU][PPT];
#pragma unroll
        for (int j = 0; j < WTA_DU; j++) {
            const int d = min(d0 + j, D - 1);
            const float *pr = row + (size_t)d * pl;
#pragma unroll
            for (int i = 0; i < PPT; i++) {
                vL[j][i] = pr[xs[i]];
                if (want_r) {
                    int xr = xs[i] + d;
                    if (xr >= W) xr -= W;
                    vR[j][i] = pr[xr];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < WTA_DU; j++) {
            const int d = d0 + j;
            if (d >= D) break;
#pragma unroll
            for (int i = 0; i < PPT; i++) {
                const float v = vL[j][i];
                if (d == 0 || (MAXIMIZE ? (v > bestL[i]) : (v < bestL[i]))) {
                    bestL[i] = v;
                    dL[i] = d;
                }
                if (want_r) {
                    const float u = vR[j][i];
                    if (d == 0 || (MAXIMIZE ? (u > bestR[i]) : (u < bestR[i]))) {
                        bestR[i] = u;
                        dR[i] = d;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < PPT; i++) {
        const int x = threadIdx.x + i * WTA_THREADS;
        if (x < W) {
            dl[rowoff + x] = (uint8_t)dL[i];
            if (beat0) beat0[rowoff + x] = dL[i] != 0;
            if (want_r) {
                dr[rowoff + x] = (uint8_t)dR[i];
                mapL[x] = (uint8_t)dL[i];
                mapR[x] = (uint8_t)dR[i];
            }
        }
    }
    if (!mask) return;
    __syncthreads();
    for (int x = threadIdx.x; x < W; x += WTA_THREADS) {
        const int v = mapL[x];
        int xt = x - v;
        if (xt < 0) xt += W;
        const int t = mapR[xt];
        mask[rowoff + x] = (uint8_t)(abs(v - t) <= tol);
    }
}

// Literal WTA carry (seq/WTA.cpp:12-15 never resets `disp`): a pixel whose
// d=0 cost is not beaten takes the disparity of the last earlier pixel in
// raster order that was beaten (0 if none).  Pass 1: per row, the last
// beaten disparity (or -1).  Pass 2 (one thread): carry into each row.
// Pass 3: per row, a sequential fill.
__global__ void k_carry_rowlast(const uint8_t *__restrict__ dl, const uint8_t *__restrict__ beat, int *__restrict__ last,
                                int W, int H) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H) return;
    int v = -1;
    for (int x = W - 1; x >= 0; x--)
        if (beat[(size_t)y * W + x]) {
            v = dl[(size_t)y * W + x];
            break;
        }
    last[y] = v;
}

__global__ void k_carry_scan(int *__restrict__ last, int H) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int carry = 0;
    for (int y = 0; y < H; y++) {
        const int v = last[y];
        last[y] = carry;   // carry into row y
        if (v >= 0) carry = v;
    }
}

__global__ void k_carry_fill(uint8_t *__restrict__ dl, const uint8_t *__restrict__ beat, const int *__restrict__ carry_in,
                             int W, int H) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= H) return;
    int c = carry_in[y];
    for (int x = 0; x < W; x++) {
        const size_t p = (size_t)y * W + x;
        if (beat[p])
            c = dl[p];
        else
            dl[p] = (uint8_t)c;
    }
}

// seq/CrossCheck.cpp:10-23 with a tolerance.
__global__ void k_cross_check(const uint8_t *__restrict__ dl, const uint8_t *__restrict__ dr, uint8_t *__restrict__ mask,
                              int W, int H, int tol) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    const int v = dl[(size_t)y * W + x];
    int xt = x - v;
    if (xt < 0) xt += W;
    const int t = dr[(size_t)y * W + xt];
    mask[(size_t)y * W + x] = (uint8_t)(abs(v - t) <= tol);
}

// seq/CrossBasedAggregation.cpp:171-212 out of place: histogram of the
// disparities in rows y-top .. y+bottom-1 (exclusive end, as the reference),
// columns x-left .. x+right of each row's own arms; most frequent, lowest on
// ties.  Histogram in registers-by-LDS: one thread per pixel with a private
// 256-bin byte-packed counter slice would not fit, so bins live in LDS per
// thread group of 64 pixels x D bins (D <= 256, counts <= 35*35).
__global__ void __launch_bounds__(64) k_disp_refine(const uint8_t *__restrict__ disp, uint8_t *__restrict__ out,
                                                    const uint32_t *__restrict__ arms, int W, int H, int D) {
    __shared__ uint16_t hist[256][64];
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y;
    for (int i = 0; i < D; i++) hist[i][threadIdx.x] = 0;
    if (x >= W) return;
    const uint32_t a = arms[(size_t)y * W + x];
    for (int row = y - arm_t(a); row < y + arm_b(a); row++) {
        if (row < 0 || row >= H) continue;
        const uint32_t ar = arms[(size_t)row * W + x];
        int xl = x - arm_l(ar), xr = x + arm_r(ar);
        if (xl < 0) xl = 0;
        if (xr >= W) xr = W - 1;
        for (int xx = xl; xx <= xr; xx++) {
            const int v = disp[(size_t)row * W + xx];
            if (v < D) hist[v][threadIdx.x]++;
        }
    }
    int best = 0, bv = hist[0][threadIdx.x];
    for (int i = 1; i < D; i++) {
        const int c = hist[i][threadIdx.x];
        if (c > bv) {
            bv = c;
            best = i;
        }
    }
    out[(size_t)y * W + x] = (uint8_t)best;
}

// Literal (in-place, raster order) refinement: inherently sequential; one
// thread walks the image.  Only for parity with the reference's in-place
// update on small inputs.
__global__ void k_disp_refine_literal(uint8_t *__restrict__ disp, const uint32_t *__restrict__ arms, int W, int H, int D,
                                      int *__restrict__ hist) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            for (int i = 0; i < D; i++) hist[i] = 0;
            const uint32_t a = arms[(size_t)y * W + x];
            for (int row = y - arm_t(a); row < y + arm_b(a); row++) {
                if (row < 0 || row >= H) continue;
                const uint32_t ar = arms[(size_t)row * W + x];
                int xl = x - arm_l(ar), xr = x + arm_r(ar);
                if (xl < 0) xl = 0;
                if (xr >= W) xr = W - 1;
                for (int xx = xl; xx <= xr; xx++) {
                    const int v = disp[(size_t)row * W + xx];
                    if (v < D) hist[v]++;
                }
            }
            int best = 0, bv = hist[0];
            for (int i = 1; i < D; i++)
                if (hist[i] > bv) {
                    bv = hist[i];
                    best = i;
                }
            disp[(size_t)y * W + x] = (uint8_t)best;
        }
}

template <bool MAXIMIZE>
static int launch_wta(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, uint8_t *beat, int W, int H, int D,
                      int tol, hipStream_t s) {
    const int ppt = div_up(W, WTA_THREADS);
    const size_t shm = 2 * (size_t)W;
#define RTS_WTA_CASE(N)                                                                         \
    case N:                                                                                     \
        k_wta<N, MAXIMIZE><<<H, WTA_THREADS, shm, s>>>(vol, dl, dr, mask, beat, W, H, D, tol); \
        break;
    switch (ppt) {
        RTS_WTA_CASE(1)
        RTS_WTA_CASE(2)
        RTS_WTA_CASE(4)
        RTS_WTA_CASE(8)
        RTS_WTA_CASE(16)
        default:
            if (ppt == 3) {
                k_wta<4, MAXIMIZE><<<H, WTA_THREADS, shm, s>>>(vol, dl, dr, mask, beat, W, H, D, tol);
            } else if (ppt <= 8) {
                k_wta<8, MAXIMIZE><<<H, WTA_THREADS, shm, s>>>(vol, dl, dr, mask, beat, W, H, D, tol);
            } else {
                k_wta<16, MAXIMIZE><<<H, WTA_THREADS, shm, s>>>(vol, dl, dr, mask, beat, W, H, D, tol);
            }
    }
#undef RTS_WTA_CASE
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // namespace rts

using namespace rts;

extern "C" {

rts_status rts_wta(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags, int tol,
                   void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(dl != nullptr, "rts_wta needs dl");
    RTS_CHECK(mask == nullptr || dr != nullptr, "the left-right mask needs dr");
    RTS_CHECK(W <= 16 * WTA_THREADS, "row width %d exceeds %d", W, 16 * WTA_THREADS);
    const bool maximize = flags & RTS_WTA_MAX;
    const bool literal = flags & RTS_WTA_LITERAL;
    hipStream_t s = S(stream);
    uint8_t *beat = nullptr;
    int *last = nullptr;
    if (literal) {
        // Scratch for the carry passes (per-pixel flags and per-row carries).
        RTS_HIP(hipMallocAsync((void **)&beat, (size_t)W * H, s));
        RTS_HIP(hipMallocAsync((void **)&last, sizeof(int) * (size_t)H, s));
    }
    int st = maximize ? launch_wta<true>(vol, dl, dr, mask, beat, W, H, D, tol, s)
                      : launch_wta<false>(vol, dl, dr, mask, beat, W, H, D, tol, s);
    if (st == RTS_OK && literal) {
        k_carry_rowlast<<<div_up(H, 64), 64, 0, s>>>(dl, beat, last, W, H);
        k_carry_scan<<<1, 1, 0, s>>>(last, H);
        k_carry_fill<<<div_up(H, 64), 64, 0, s>>>(dl, beat, last, W, H);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = fail(RTS_ERR_HIP, "carry launch: %s", hipGetErrorString(e));
        // The literal carry changes D_L after the fused check ran: redo it.
        if (st == RTS_OK && mask) {
            k_cross_check<<<dim3(div_up(W, 256), H), 256, 0, s>>>(dl, dr, mask, W, H, tol);
            e = hipGetLastError();
            if (e != hipSuccess) st = fail(RTS_ERR_HIP, "cross-check launch: %s", hipGetErrorString(e));
        }
    }
    if (beat) (void)hipFreeAsync(beat, s);
    if (last) (void)hipFreeAsync(last, s);
    return st;
}

rts_status rts_cross_check(const uint8_t *dl, const uint8_t *dr, uint8_t *mask, int W, int H, int tol, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && H <= 65535, "bad image size %dx%d", W, H);
    k_cross_check<<<dim3(div_up(W, 256), H), 256, 0, S(stream)>>>(dl, dr, mask, W, H, tol);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_disp_refine(uint8_t *disp, uint8_t *out, const uint32_t *arms, int W, int H, int D, int literal,
                           void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    hipStream_t s = S(stream);
    if (literal) {
        int *hist = nullptr;
        RTS_HIP(hipMallocAsync((void **)&hist, sizeof(int) * 256, s));
        k_disp_refine_literal<<<1, 1, 0, s>>>(disp, arms, W, H, D, hist);
        hipError_t e = hipGetLastError();
        (void)hipFreeAsync(hist, s);
        if (e != hipSuccess) return fail(RTS_ERR_HIP, "refine launch: %s", hipGetErrorString(e));
        return RTS_OK;
    }
    RTS_CHECK(out != nullptr && out != disp, "out-of-place refinement needs a separate output");
    RTS_CHECK(H <= 65535, "height %d too large", H);
    k_disp_refine<<<dim3(div_up(W, 64), H), 64, 0, s>>>(disp, out, arms, W, H, D);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
