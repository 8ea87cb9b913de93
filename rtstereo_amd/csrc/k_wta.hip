// Disparity selection: winner-takes-all with the right view read off the
// volume's diagonal, the left-right check, and the cross-support histogram
// refinement.
#include "rts_internal.h"

namespace rts {

// A block of T threads owns a chunk of T*PPT pixels of one row; thread t
// owns pixels t, t+T, ...  For every d it reads vol[d][y][x] for the left view and
// vol[d][y][(x+d) mod W] for the right view (right pixel x matched with left
// pixel x+d); both reads are contiguous across the wavefront, and the second
// mostly hits the cache lines the first brought in.  d = 0 seeds both
// searches and strict comparisons from d = 1 up keep the lowest d on ties
// (seq/WTA.cpp:16-21).  DU disparities are loaded before they are compared,
// so their loads are in flight together.  The left-right check
// (seq/CrossCheck.cpp:15-21) is k_cross_check on the finished maps.

template <typename T>
__device__ __forceinline__ T at_byte(const T *base, uint32_t off) {
    return *(const T *)((const char *)base + off);
}

// A value the compiler cannot see through, so that a 32-bit offset stays a
// 32-bit per-lane operand of a scalar-base load.
__device__ __forceinline__ uint32_t opaque_u32(uint32_t v) {
    asm volatile("" : "+v"(v));
    return v;
}

template <bool MAXIMIZE>
__device__ __forceinline__ bool beats(float v, float best) {
    return MAXIMIZE ? v > best : v < best;
}

template <int PPT, bool MAXIMIZE, bool WANT_R, int DU, int T>
__global__ void __launch_bounds__(T) k_wta(const float *__restrict__ vol, uint8_t *__restrict__ dl,
                                                     uint8_t *__restrict__ dr, uint8_t *__restrict__ beat0, int W,
                                                     int H, int D) {
    // Chunks of a row are consecutive in the work order and each XCD takes
    // one contiguous range of it, so the right view's reads past the chunk
    // (columns up to x + D - 1) find the next chunk's lines in the same L2.
    const int nch = gridDim.x;
    const int wk = xcd_order(blockIdx.y * nch + blockIdx.x, nch * gridDim.y);
    const int y = wk / nch;
    const int xb = (wk - y * nch) * (T * PPT) + threadIdx.x;
    const size_t pl = plane(W, H), rowoff = (size_t)y * W;
    const float *row = vol + rowoff;
    // byte offsets of the thread's pixels, and of the same pixels minus a row
    // (the right view reads column (x + d) mod W = min(x+d, x+d-W) unsigned)
    uint32_t ol[PPT], om[PPT];
    float bestL[PPT], bestR[PPT];
    int dL[PPT], dR[PPT];
#pragma unroll
    for (int i = 0; i < PPT; i++) {
        const uint32_t x = (uint32_t)min(xb + i * T, W - 1);
        ol[i] = x * 4u;
        om[i] = (x - (uint32_t)W) * 4u;
        bestL[i] = at_byte(row, ol[i]);   // d = 0 seeds the search
        bestR[i] = bestL[i];
        dL[i] = dR[i] = 0;
    }
    for (int d0 = 1; d0 < D; d0 += DU) {
        float vL[DU][PPT], vR[DU][PPT];
#pragma unroll
        for (int j = 0; j < DU; j++) {
            const int d = min(d0 + j, D - 1);
            const float *pr = row + (size_t)d * pl;
            const uint32_t d4 = (uint32_t)d * 4u;
#pragma unroll
            for (int i = 0; i < PPT; i++) {
                vL[j][i] = at_byte(pr, opaque_u32(ol[i]));
                if (WANT_R) vR[j][i] = at_byte(pr, min(ol[i] + d4, om[i] + d4));
            }
        }
        // every load of the batch is issued before the first comparison waits
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int j = 0; j < DU; j++) {
            const int d = d0 + j;
            if (d >= D) break;
#pragma unroll
            for (int i = 0; i < PPT; i++) {
                if (beats<MAXIMIZE>(vL[j][i], bestL[i])) {
                    bestL[i] = vL[j][i];
                    dL[i] = d;
                }
                if (WANT_R && beats<MAXIMIZE>(vR[j][i], bestR[i])) {
                    bestR[i] = vR[j][i];
                    dR[i] = d;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < PPT; i++) {
        const int x = xb + i * T;
        if (x < W) {
            dl[rowoff + x] = (uint8_t)dL[i];
            if (beat0) beat0[rowoff + x] = dL[i] != 0;
            if (WANT_R) dr[rowoff + x] = (uint8_t)dR[i];
        }
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

// Rows are cut into chunks, one block each.  One wave per block and 8 pixels per lane (512-pixel row chunks): a 1080p
// frame gives 4320 blocks, all resident at once, spread evenly over the CUs.
// With 8 pixels per lane one disparity at a time (2*8 loads in flight per
// lane) measured fastest (0.27 against 0.29 ms with four at a time at
// 1080p/192); narrower rows, with fewer pixels per lane, batch four.
constexpr int WTA_T = 64;

template <int PPT, bool MAXIMIZE, bool WANT_R>
static void wta_go(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *beat, int W, int H, int D, hipStream_t s) {
    constexpr int DU = PPT >= 8 ? 1 : 4;
    k_wta<PPT, MAXIMIZE, WANT_R, DU, WTA_T><<<dim3(div_up(W, WTA_T * PPT), H), WTA_T, 0, s>>>(vol, dl, dr, beat, W, H,
                                                                                             D);
}

template <bool MAXIMIZE, bool WANT_R>
static void launch_wta_r(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *beat, int W, int H, int D,
                         hipStream_t s) {
    if (W <= 64)
        wta_go<1, MAXIMIZE, WANT_R>(vol, dl, dr, beat, W, H, D, s);
    else if (W <= 128)
        wta_go<2, MAXIMIZE, WANT_R>(vol, dl, dr, beat, W, H, D, s);
    else if (W <= 256)
        wta_go<4, MAXIMIZE, WANT_R>(vol, dl, dr, beat, W, H, D, s);
    else
        wta_go<8, MAXIMIZE, WANT_R>(vol, dl, dr, beat, W, H, D, s);
}

template <bool MAXIMIZE>
static int launch_wta(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *beat, int W, int H, int D,
                      hipStream_t s) {
    RTS_CHECK(H <= 65535, "image height %d too large", H);
    if (dr)
        launch_wta_r<MAXIMIZE, true>(vol, dl, dr, beat, W, H, D, s);
    else
        launch_wta_r<MAXIMIZE, false>(vol, dl, dr, beat, W, H, D, s);
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
    int st = maximize ? launch_wta<true>(vol, dl, dr, beat, W, H, D, s) : launch_wta<false>(vol, dl, dr, beat, W, H, D, s);
    if (st == RTS_OK && literal) {
        k_carry_rowlast<<<div_up(H, 64), 64, 0, s>>>(dl, beat, last, W, H);
        k_carry_scan<<<1, 1, 0, s>>>(last, H);
        k_carry_fill<<<div_up(H, 64), 64, 0, s>>>(dl, beat, last, W, H);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = fail(RTS_ERR_HIP, "carry launch: %s", hipGetErrorString(e));
    }
    // the left-right check runs on the final D_L (after the literal carry)
    if (st == RTS_OK && mask) {
        k_cross_check<<<dim3(div_up(W, 256), H), 256, 0, s>>>(dl, dr, mask, W, H, tol);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) st = fail(RTS_ERR_HIP, "cross-check launch: %s", hipGetErrorString(e));
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
