// Disparity selection: winner-takes-all with the right view read off the
// volume's diagonal, the left-right check, and the cross-support histogram
// refinement.
#include "rts_internal.h"

#include <mutex>

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
                                                     uint8_t *__restrict__ dr, int W, int H, int D) {
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
            if (WANT_R) dr[rowoff + x] = (uint8_t)dR[i];
        }
    }
}

// Literal WTA carry (seq/WTA.cpp:12-15 never resets `disp`): a pixel keeps
// the previous pixel's disparity (raster order) unless some d > 0 beats the
// running best, which starts at d = 0.  With the lowest d winning ties, a
// pixel is beaten exactly when its own argbest is not 0, so the literal map
// is: the argbest where it is not 0, else the last non-zero argbest before
// the pixel in raster order (0 if there is none).  Three passes over the
// finished map, with one int of scratch per row:
//   1. per row, the last non-zero disparity (-1 if none);
//   2. one block turns those into the carry into each row (an exclusive
//      "latest non-negative" scan over rows);
//   3. per row, zeros take the carry, or the last non-zero left of them.
// Rows are one wave each; lane j walks the row segment [j*c, (j+1)*c).
__device__ __forceinline__ int seg_last_nonzero(const uint8_t *row, int x0, int x1) {
    int v = -1;
    for (int x = x0; x < x1; x++) {
        const int t = row[x];
        if (t) v = t;
    }
    return v;
}

__global__ void __launch_bounds__(64) k_carry_rowlast(const uint8_t *__restrict__ dl, int *__restrict__ last, int W) {
    const int y = blockIdx.x, lane = threadIdx.x;
    const int c = (W + 63) / 64, x0 = min(lane * c, W), x1 = min(x0 + c, W);
    const int v = seg_last_nonzero(dl + (size_t)y * W, x0, x1);
    const uint64_t m = __ballot(v >= 0);
    const int src = m ? 63 - __clzll((long long)m) : 0;
    const int r = __shfl(v, src);
    if (lane == 0) last[y] = m ? r : -1;
}

// carry[y] <- the last last[j] >= 0 with j < y, else 0 (in place).
__global__ void __launch_bounds__(1024) k_carry_scan(int *__restrict__ carry, int H) {
    __shared__ int key[1024], val[1024];
    const int t = threadIdx.x;
    int in = 0;   // carried into the current chunk of rows
    for (int base = 0; base < H; base += 1024) {
        const int i = base + t;
        const int v = i < H ? carry[i] : -1;
        val[t] = v;
        key[t] = v >= 0 ? t : -1;
        __syncthreads();
        // inclusive max-scan of key: the latest row <= t of the chunk with a value
        for (int off = 1; off < 1024; off <<= 1) {
            const int k = t >= off ? key[t - off] : -1;
            __syncthreads();
            if (k > key[t]) key[t] = k;
            __syncthreads();
        }
        const int kp = t > 0 ? key[t - 1] : -1;
        if (i < H) carry[i] = kp >= 0 ? val[kp] : in;
        const int kl = key[1023];
        in = kl >= 0 ? val[kl] : in;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(64) k_carry_fill(uint8_t *__restrict__ dl, const int *__restrict__ carry, int W) {
    const int y = blockIdx.x, lane = threadIdx.x;
    const int c = (W + 63) / 64, x0 = min(lane * c, W), x1 = min(x0 + c, W);
    uint8_t *row = dl + (size_t)y * W;
    const int v = seg_last_nonzero(row, x0, x1);
    // the latest lower lane with a non-zero disparity, else the row's carry
    const uint64_t m = __ballot(v >= 0) & ((1ull << lane) - 1ull);
    const int src = m ? 63 - __clzll((long long)m) : 0;
    const int r = __shfl(v, src);
    int cur = m ? r : carry[y];
    for (int x = x0; x < x1; x++) {
        const int t = row[x];
        if (t)
            cur = t;
        else
            row[x] = (uint8_t)cur;
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
// ties.  One thread per pixel, 64 pixels per block, 16-bit bins in LDS
// ([256][64]).  A support of more than 65535 pixels (only possible with arms
// above 127) could wrap a 16-bit bin, so such a pixel counts again with a
// 32-bit counter per disparity value.
__device__ int refine_mode_wide(const uint8_t *__restrict__ disp, const uint32_t *__restrict__ arms, int x, int y,
                                int W, int H, int D) {
    const uint32_t a = arms[(size_t)y * W + x];
    int best = 0;
    uint32_t bv = 0;
    for (int v = 0; v < D; v++) {
        uint32_t c = 0;
        for (int row = max(y - arm_t(a), 0); row < min(y + arm_b(a), H); row++) {
            const uint32_t ar = arms[(size_t)row * W + x];
            const int xl = max(x - arm_l(ar), 0), xr = min(x + arm_r(ar), W - 1);
            for (int xx = xl; xx <= xr; xx++) c += disp[(size_t)row * W + xx] == v;
        }
        if (c > bv || v == 0) {
            bv = c;
            best = v;
        }
    }
    return best;
}

__global__ void __launch_bounds__(64) k_disp_refine(const uint8_t *__restrict__ disp, uint8_t *__restrict__ out,
                                                    const uint32_t *__restrict__ arms, int W, int H, int D) {
    __shared__ uint16_t hist[256][64];
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y;
    for (int i = 0; i < D; i++) hist[i][threadIdx.x] = 0;
    if (x >= W) return;
    const uint32_t a = arms[(size_t)y * W + x];
    uint32_t tot = 0;
    for (int row = y - arm_t(a); row < y + arm_b(a); row++) {
        if (row < 0 || row >= H) continue;
        const uint32_t ar = arms[(size_t)row * W + x];
        int xl = x - arm_l(ar), xr = x + arm_r(ar);
        if (xl < 0) xl = 0;
        if (xr >= W) xr = W - 1;
        tot += (uint32_t)(xr - xl + 1);
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
    if (tot > 65535u) best = refine_mode_wide(disp, arms, x, y, W, H, D);
    out[(size_t)y * W + x] = (uint8_t)best;
}

// Literal refinement: the reference updates dispIm in place in raster order
// (seq/CrossBasedAggregation.cpp:175-211), so a pixel's vote sees the refined
// values of the pixels before it and the original values of those after it.
// Pixel (y, x) reads only columns within A = the largest left or right arm of
// x.  With K = A + 1, every pixel it reads in an earlier row lies on an
// earlier anti-diagonal s = x + K*y, every pixel in a later row on a later
// one, and two pixels of one diagonal on different rows are at least K
// columns apart, so they never read each other.  One block therefore walks
// the diagonals in order with a barrier between them; its waves refine a
// diagonal's pixels in parallel, one wave per pixel with the histogram in
// LDS (32-bit bins).  A wave first loads the arms of all its support rows
// (lane = row), then the rows' disparities eight rows at a time (lanes over
// columns), so that their loads are in flight together.
constexpr int RF_WAVES = 16, RF_UNROLL = 8;

__global__ void __launch_bounds__(64 * RF_WAVES) k_disp_refine_diag(uint8_t *__restrict__ disp,
                                                                    const uint32_t *__restrict__ arms, int W, int H,
                                                                    int D) {
    __shared__ uint32_t hist[RF_WAVES][256];
    __shared__ uint32_t wmax[RF_WAVES];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // A: the largest horizontal arm of the image
    uint32_t am = 0;
    for (size_t i = threadIdx.x; i < (size_t)W * H; i += blockDim.x) {
        const uint32_t a = arms[i];
        am = max(am, max(a & 255u, (a >> 8) & 255u));
    }
    for (int off = 32; off >= 1; off >>= 1) am = max(am, (uint32_t)__shfl_xor((int)am, off));
    if (lane == 0) wmax[w] = am;
    __syncthreads();
    for (int i = 0; i < RF_WAVES; i++) am = max(am, wmax[i]);
    const long long K = (long long)am + 1;
    const long long T = (long long)(W - 1) + K * (H - 1);
    uint32_t *hw = hist[w];
    for (long long t = 0; t <= T; t++) {
        const int ylo = t >= W ? (int)((t - W + K) / K) : 0;
        const int yhi = (int)min((long long)(H - 1), t / K);
        for (int y = ylo + w; y <= yhi; y += RF_WAVES) {
            const int x = (int)(t - K * y);
            for (int i = lane; i < 256; i += 64) hw[i] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            const uint32_t a = arms[(size_t)y * W + x];
            const int r0 = max(y - arm_t(a), 0), r1 = min(y + arm_b(a), H);
            for (int jb = r0; jb < r1; jb += 64) {
                const int nb = min(64, r1 - jb);
                // lane j: the extent of support row jb + j
                int xl = 0, n = 0;
                if (lane < nb) {
                    const uint32_t ar = arms[(size_t)(jb + lane) * W + x];
                    xl = max(x - arm_l(ar), 0);
                    n = min(x + arm_r(ar), W - 1) - xl + 1;
                }
                for (int k0 = 0; k0 < nb; k0 += RF_UNROLL) {
                    int v[RF_UNROLL];
#pragma unroll
                    for (int u = 0; u < RF_UNROLL; u++) {
                        const int k = min(k0 + u, nb - 1);
                        const int rxl = __shfl(xl, k), rn = __shfl(n, k);
                        v[u] = k0 + u < nb && lane < rn ? disp[(size_t)(jb + k) * W + rxl + lane] : 256;
                    }
#pragma unroll
                    for (int u = 0; u < RF_UNROLL; u++)
                        if (v[u] < D) atomicAdd(&hw[v[u]], 1u);
                    // rows wider than 64 columns (arms above 31)
                    for (int u = 0; u < RF_UNROLL && k0 + u < nb; u++) {
                        const int k = k0 + u;
                        const int rxl = __shfl(xl, k), rn = __shfl(n, k);
                        for (int c = 64 + lane; c < rn; c += 64) {
                            const int vv = disp[(size_t)(jb + k) * W + rxl + c];
                            if (vv < D) atomicAdd(&hw[vv], 1u);
                        }
                    }
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // most frequent value, lowest on ties: each lane its four bins in
            // ascending order, then the wave's best (count, then index)
            uint32_t bc = 0;
            int bi = 256;
            for (int i = lane; i < D; i += 64)
                if (bi == 256 || hw[i] > bc) {
                    bc = hw[i];
                    bi = i;
                }
            for (int off = 32; off >= 1; off >>= 1) {
                const uint32_t oc = (uint32_t)__shfl_xor((int)bc, off);
                const int oi = __shfl_xor(bi, off);
                if (oi != 256 && (bi == 256 || oc > bc || (oc == bc && oi < bi))) {
                    bc = oc;
                    bi = oi;
                }
            }
            if (lane == 0) disp[(size_t)y * W + x] = (uint8_t)bi;
        }
        __syncthreads();
    }
}

// Rows are cut into chunks, one block each.  One wave per block and 8 pixels per lane (512-pixel row chunks): a 1080p
// frame gives 4320 blocks, all resident at once, spread evenly over the CUs.
// With 8 pixels per lane one disparity at a time (2*8 loads in flight per
// lane) measured fastest (0.27 against 0.29 ms with four at a time at
// 1080p/192); narrower rows, with fewer pixels per lane, batch four.
constexpr int WTA_T = 64;

template <int PPT, bool MAXIMIZE, bool WANT_R>
static void wta_go(const float *vol, uint8_t *dl, uint8_t *dr, int W, int H, int D, hipStream_t s) {
    constexpr int DU = PPT >= 8 ? 1 : 4;
    k_wta<PPT, MAXIMIZE, WANT_R, DU, WTA_T><<<dim3(div_up(W, WTA_T * PPT), H), WTA_T, 0, s>>>(vol, dl, dr, W, H, D);
}

template <bool MAXIMIZE, bool WANT_R>
static void launch_wta_r(const float *vol, uint8_t *dl, uint8_t *dr, int W, int H, int D, hipStream_t s) {
    if (W <= 64)
        wta_go<1, MAXIMIZE, WANT_R>(vol, dl, dr, W, H, D, s);
    else if (W <= 128)
        wta_go<2, MAXIMIZE, WANT_R>(vol, dl, dr, W, H, D, s);
    else if (W <= 256)
        wta_go<4, MAXIMIZE, WANT_R>(vol, dl, dr, W, H, D, s);
    else
        wta_go<8, MAXIMIZE, WANT_R>(vol, dl, dr, W, H, D, s);
}

template <bool MAXIMIZE>
static void launch_wta(const float *vol, uint8_t *dl, uint8_t *dr, int W, int H, int D, hipStream_t s) {
    if (dr)
        launch_wta_r<MAXIMIZE, true>(vol, dl, dr, W, H, D, s);
    else
        launch_wta_r<MAXIMIZE, false>(vol, dl, dr, W, H, D, s);
}

int wta_launch(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags, int tol,
               int *carry, hipStream_t s) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(dl != nullptr, "rts_wta needs dl");
    RTS_CHECK(mask == nullptr || dr != nullptr, "the left-right mask needs dr");
    RTS_CHECK(H <= WTA_MAX_ROWS, "image height %d too large (at most %d rows)", H, WTA_MAX_ROWS);
    const bool literal = flags & RTS_WTA_LITERAL;
    RTS_CHECK(!literal || carry != nullptr, "the literal carry needs its per-row scratch");
    if (flags & RTS_WTA_MAX)
        launch_wta<true>(vol, dl, dr, W, H, D, s);
    else
        launch_wta<false>(vol, dl, dr, W, H, D, s);
    RTS_LAUNCHED();
    if (literal) {
        k_carry_rowlast<<<H, 64, 0, s>>>(dl, carry, W);
        k_carry_scan<<<1, 1024, 0, s>>>(carry, H);
        k_carry_fill<<<H, 64, 0, s>>>(dl, carry, W);
        RTS_LAUNCHED();
    }
    // the left-right check runs on the final D_L (after the literal carry)
    if (mask) {
        k_cross_check<<<dim3(div_up(W, 256), H), 256, 0, s>>>(dl, dr, mask, W, H, tol);
        RTS_LAUNCHED();
    }
    return RTS_OK;
}

// Scratch of rts_wta_noalloc: the per-row carries of the literal WTA.  Each
// device holds its own copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress returns for the current
// device (the array's name is not a device address on the host).
static __device__ __align__(256) int scratch[WTA_MAX_ROWS];

namespace {
std::mutex g_scratch_mu;
int *g_scratch_ptr[RTS_MAX_DEVICES] = {};

// Device address of `scratch` on the current device, looked up on the
// device's first call.
int scratch_ptr(int **out) {
    int dev = 0;
    RTS_HIP(hipGetDevice(&dev));
    RTS_CHECK(dev >= 0 && dev < RTS_MAX_DEVICES, "device %d beyond the %d supported", dev, RTS_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        RTS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RTS_CHECK(p != nullptr, "no device address for the WTA scratch");
        g_scratch_ptr[dev] = (int *)p;
    }
    *out = g_scratch_ptr[dev];
    return RTS_OK;
}
}  // namespace

}  // namespace rts

using namespace rts;

extern "C" {

rts_status rts_wta(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags, int tol,
                   void *stream) {
    hipStream_t s = S(stream);
    int *carry = nullptr;
    if ((flags & RTS_WTA_LITERAL) && H >= 1 && H <= WTA_MAX_ROWS)
        RTS_HIP(hipMallocAsync((void **)&carry, sizeof(int) * (size_t)H, s));
    const int st = wta_launch(vol, dl, dr, mask, W, H, D, flags, tol, carry, s);
    if (carry) (void)hipFreeAsync(carry, s);
    return st;
}

rts_status rts_wta_noalloc(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags,
                           int tol, void *stream) {
    RTS_CHECK(H >= 1 && H <= WTA_MAX_ROWS, "rts_wta_noalloc: height %d outside 1..%d (its scratch holds %d rows)", H,
              WTA_MAX_ROWS, WTA_MAX_ROWS);
    int *carry = nullptr;
    if (int e = scratch_ptr(&carry)) return e;
    return wta_launch(vol, dl, dr, mask, W, H, D, flags, tol, carry, S(stream));
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
        k_disp_refine_diag<<<1, 64 * RF_WAVES, 0, s>>>(disp, arms, W, H, D);
        RTS_LAUNCHED();
        return RTS_OK;
    }
    RTS_CHECK(out != nullptr && out != disp, "out-of-place refinement needs a separate output");
    RTS_CHECK(H <= 65535, "height %d too large", H);
    k_disp_refine<<<dim3(div_up(W, 64), H), 64, 0, s>>>(disp, out, arms, W, H, D);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
