// Cost aggregation kernels: box mean, adaptive weights, cross arms, the
// serial row prefix and the vertical cross sweep.  All volumes are
// disparity-major [D][H][W].
#include "rts_internal.h"

#include <stdlib.h>

#include <algorithm>
#include <type_traits>

#pragma clang fp contract(off)

namespace rts {

// ---------------------------------------------------------------- box

// seq/ConstBlockAggregation.cpp:11-34, out of place.  Every output sums its
// (2r+1)^2 window in the reference's row-major order (rows outer, columns
// inner, one float accumulator) and divides by the window area.

// Radii 0..BOX_RT: one block = one d plane tile of 64 x BOXR_TY outputs.  The
// (64+2R) x (BOXR_TY+2R) input tile is staged in LDS; a thread owns one
// column and BOXR_RPT consecutive rows, reads each input row's 2R+1 taps
// once and adds every tap to the outputs whose window holds that row, so a
// tap costs one LDS read per BOXR_RPT outputs instead of one per output.
// Each accumulator still receives its taps rows-ascending, columns-ascending.
constexpr int BOXR_TY = 32;
constexpr int BOXR_RPT = BOXR_TY / 4;
constexpr int BOX_RT = 8;

template <int R>
__global__ void __launch_bounds__(256) k_box_rw(const float *__restrict__ in, float *__restrict__ out, int W, int H) {
    constexpr int E = 2 * R + 1, TH = BOXR_TY + 2 * R, TW = 64 + 2 * R;
    __shared__ float tile[TH][TW];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * BOXR_TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const float *src = in + base;
    const bool inner = x0 - R >= 0 && x0 + 64 + R <= W && y0 - R >= 0 && y0 + BOXR_TY + R <= H;
    // All of a thread's loads are issued before its LDS stores.
    constexpr int NB = (TH * TW + 255) / 256;
    float v[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int i = threadIdx.x + 256 * k;
        const int rr = i / TW, cc = i - rr * TW;
        float x = 0.0f;
        if (rr < TH) {
            const int yy = y0 - R + rr, xx = x0 - R + cc;
            x = inner ? src[(size_t)yy * W + xx] : src[(size_t)wrapn(yy, H) * W + wrapn(xx, W)];
        }
        v[k] = x;
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int i = threadIdx.x + 256 * k;
        const int rr = i / TW, cc = i - rr * TW;
        if (rr < TH) tile[rr][cc] = v[k];
    }
    __syncthreads();
    const int tx = threadIdx.x & 63, r0 = (threadIdx.x >> 6) * BOXR_RPT;
    float acc[BOXR_RPT];
#pragma unroll
    for (int i = 0; i < BOXR_RPT; i++) acc[i] = 0.0f;
#pragma unroll
    for (int rr = 0; rr < BOXR_RPT + 2 * R; rr++) {
#pragma unroll
        for (int dx = 0; dx < E; dx++) {
            const float t = tile[r0 + rr][tx + dx];
#pragma unroll
            for (int i = 0; i < BOXR_RPT; i++)
                if (rr >= i && rr < i + E) acc[i] += t;
        }
    }
    const int x = x0 + tx;
    if (x >= W) return;
#pragma unroll
    for (int i = 0; i < BOXR_RPT; i++) {
        const int y = y0 + r0 + i;
        if (y < H) stream_store(&out[base + (size_t)y * W + x], div_count(acc[i], (uint32_t)(E * E)));
    }
}

// Radii BOX_RT+1..BOX_RMAX: one block = one d plane tile of 64 x BOX_TY
// outputs, every tap read from the LDS tile.
constexpr int BOX_TY = 16;
constexpr int BOX_RMAX = 16;

__global__ void __launch_bounds__(256) k_box_f32(const float *__restrict__ in, float *__restrict__ out, int W, int H,
                                                 int r) {
    __shared__ float tile[BOX_TY + 2 * BOX_RMAX][64 + 2 * BOX_RMAX + 1];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * BOX_TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const int tw = 64 + 2 * r, th = BOX_TY + 2 * r;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    // All of a thread's loads are issued before its LDS stores.
    constexpr int NB = ((BOX_TY + 2 * BOX_RMAX) * (64 + 2 * BOX_RMAX) + 255) / 256;
    float v[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int i = tid + 256 * k;
        const int ty = i / tw, tx = i - ty * tw;
        v[k] = ty < th ? in[base + (size_t)wrapn(y0 - r + ty, H) * W + wrapn(x0 - r + tx, W)] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int i = tid + 256 * k;
        const int ty = i / tw, tx = i - ty * tw;
        if (ty < th) tile[ty][tx] = v[k];
    }
    __syncthreads();
    const float area = (float)((2 * r + 1) * (2 * r + 1));
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
    for (int ly = threadIdx.y; ly < BOX_TY; ly += 4) {
        const int y = y0 + ly;
        if (y >= H) break;
        float acc = 0.0f;
        for (int dy = 0; dy <= 2 * r; dy++)
            for (int dx = 0; dx <= 2 * r; dx++) acc += tile[ly + dy][threadIdx.x + dx];
        stream_store(&out[base + (size_t)y * W + x], __fdiv_rn(acc, area));
    }
}

// Fallback for radii above BOX_RMAX: window read through the caches.
__global__ void __launch_bounds__(256) k_box_f32_global(const float *__restrict__ in, float *__restrict__ out, int W,
                                                        int H, int r) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y, d = blockIdx.z;
    if (x >= W || y >= H) return;
    const size_t base = (size_t)d * plane(W, H);
    float acc = 0.0f;
    for (int dy = -r; dy <= r; dy++) {
        const size_t rr = base + (size_t)wrap1(y + dy, H) * W;
        for (int dx = -r; dx <= r; dx++) acc += in[rr + wrap1(x + dx, W)];
    }
    out[base + (size_t)y * W + x] = __fdiv_rn(acc, (float)((2 * r + 1) * (2 * r + 1)));
}

// Box over integer-valued costs by running sums (seq/ConstBlockAggregation.cpp:11-34
// for volumes whose every value is an integer in [0, 2^24 / (2r+1)^2]: AD of
// bytes, truncated AD at an integer threshold, census Hamming, rank AD).
// Every partial sum of such a window is an exact integer, so the float sum
// the reference forms in row-major order equals the integer window sum, and
// the quotient by the area is the same float.
//
// A wave owns one d plane, an aligned tile of 64 output columns x0 .. x0+63
// and a segment of rows.  Per input row it loads the tile's 64 columns and,
// on lanes 0 .. r-1 and 64-r .. 63, the 2r halo columns beside it (wrapped),
// forms the inclusive prefix of each across the wave (DPP row shifts, then
// the row broadcasts), and takes each output's row sum h from prefix values
// (ds_bpermute): the tile part S[min(j+r,63)] - S[j-r-1], plus the part of
// the left halo at or right of column j-r, plus the part of the right halo up
// to column j+r.  The column sum runs down the segment:
// V(y+1) = V(y) + h(y+r+1) - h(y-r), the last 2r+1 row sums kept in a
// per-wave LDS ring.  O(1) adds per output instead of (2r+1)^2; loads and
// stores are whole aligned rows of the tile.  Waves are independent (no
// block barrier).
constexpr int BOXI_RMAX = 16;
constexpr int BOXI_PF = 8;   // input rows in flight per wave

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);   // row_bcast:15 into rows 1, 3
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);   // row_bcast:31 into rows 2, 3
    return v;
}

__global__ void __launch_bounds__(256) k_box_int(const float *__restrict__ in, float *__restrict__ out, int W, int H,
                                                 int r, int ntx, int nseg, int seg, int nwork) {
    extern __shared__ uint32_t boxi_ring[];   // per wave: 2r+1 rows of 64 row sums
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int work = blockIdx.x * 4 + wv;
    if (work >= nwork) return;   // the whole wave leaves
    const int xt = work % ntx, rest = work / ntx;
    const int sg = rest % nseg, d = rest / nseg;
    const int E = 2 * r + 1;
    const int x0 = xt * 64, x = x0 + lane;
    const bool outlane = x < W;
    const int cmain = wrapn(x, W);
    // halo lanes: j < r holds column x0-r+j, j >= 64-r holds column x0+j+r
    const bool lhalo = lane < r, rhalo = lane >= 64 - r;
    const int chalo = lhalo ? wrapn(x0 - r + lane, W) : rhalo ? wrapn(x0 + lane + r, W) : cmain;
    const int s0 = sg * seg, s1 = min(H, s0 + seg);
    const size_t base = (size_t)d * plane(W, H);
    const float *src = in + base;
    float *dst = out + base;
    uint32_t(*rg)[64] = reinterpret_cast<uint32_t(*)[64]>(boxi_ring + (size_t)wv * E * 64);
    const int a_hi = min(lane + r, 63) * 4, a_lo = ((lane - r - 1) & 63) * 4, a_prev = ((lane - 1) & 63) * 4;
    const bool has_lo = lane > r;
    auto load = [&](int y) {
        const float *row = src + (size_t)wrapn(y, H) * W;
        return make_float2(row[cmain], (lhalo || rhalo) ? row[chalo] : 0.0f);
    };
    // window sum over the output's 2r+1 columns of one input row
    auto hrow = [&](float2 v) -> uint32_t {
        const uint32_t S = wave_incl_scan((uint32_t)v.x);
        const uint32_t U = wave_incl_scan((uint32_t)v.y);   // zero between the two halos
        const uint32_t hi = (uint32_t)__builtin_amdgcn_ds_bpermute(a_hi, (int)S);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_ds_bpermute(a_lo, (int)S);
        const uint32_t up = (uint32_t)__builtin_amdgcn_ds_bpermute(a_prev, (int)U);
        const uint32_t ul = r > 0 ? (uint32_t)__builtin_amdgcn_readlane((int)U, r - 1) : 0u;   // the left halo's sum
        uint32_t h = hi - (has_lo ? lo : 0u);
        if (lhalo) h += ul - (lane > 0 ? up : 0u);   // left halo columns j-r .. x0-1
        if (rhalo) h += U - ul;                       // right halo columns x0+64 .. j+r
        return h;
    };
    const uint32_t area = (uint32_t)(E * E);
    // the first window: rows s0-r .. s0+r into ring slots 0 .. 2r
    uint32_t V = 0;
    for (int i0 = 0; i0 < E; i0 += 4) {
        float2 v[4];
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = load(s0 - r + min(i0 + k, E - 1));
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (i0 + k < E) {
                const uint32_t h = hrow(v[k]);
                rg[i0 + k][lane] = h;
                V += h;
            }
    }
    // q[k] holds the input row entering the window at step y = yb + k
    float2 q[BOXI_PF];
#pragma unroll
    for (int k = 0; k < BOXI_PF; k++) q[k] = load(s0 + r + 1 + k);
    int slot = 0;   // ring slot of row y - r
    auto step = [&](int y, uint32_t hn) {
        if (outlane) stream_store(&dst[(size_t)y * W + x], div_count((float)V, area));
        V += hn - rg[slot][lane];
        rg[slot][lane] = hn;
        if (++slot == E) slot = 0;
    };
    int yb = s0;
    // whole groups of BOXI_PF rows: the group's row sums are independent of
    // each other and formed first, then the column sum steps through them
    for (; yb + BOXI_PF <= s1; yb += BOXI_PF) {
        uint32_t hn[BOXI_PF];
#pragma unroll
        for (int k = 0; k < BOXI_PF; k++) hn[k] = hrow(q[k]);   // row yb + k + r + 1
#pragma unroll
        for (int k = 0; k < BOXI_PF; k++) q[k] = load(yb + k + r + 1 + BOXI_PF);
#pragma unroll
        for (int k = 0; k < BOXI_PF; k++) step(yb + k, hn[k]);
    }
#pragma unroll
    for (int k = 0; k < BOXI_PF; k++)
        if (yb + k < s1) step(yb + k, hrow(q[k]));
}

// ---------------------------------------------------------------- adaptive weight

// seq/AdaptiveWeightAggregation.cpp:11-18.
__global__ void k_prox(float *prox, float gp, int r) {
    const int e = 2 * r + 1;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= e * e) return;
    const int dy = k / e - r, dx = k % e - r;
    prox[k] = expf(__fdiv_rn(-sqrt_rn((float)(dy * dy + dx * dx)), gp));
}

// Lab distance in the reference's float order (seq/AdaptiveWeightAggregation.cpp:19-24).
__device__ __forceinline__ float lab_dist(const uint8_t *a, const uint8_t *b) {
    const float e0 = (float)a[0] - (float)b[0], e1 = (float)a[1] - (float)b[1], e2 = (float)a[2] - (float)b[2];
    return sqrt_rn(e0 * e0 + e1 * e1 + e2 * e2);
}

// seq/AdaptiveWeightAggregation.cpp:33-71: w[p][k] = prox[k]*exp(-dist/gc),
// 0 outside the image.  One thread per (pixel, k).
__global__ void k_weights(float *__restrict__ w, const float *__restrict__ prox, const uint8_t *__restrict__ lab,
                          float gc, int W, int H, int r) {
    const int e = 2 * r + 1, A = e * e;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)A * W * H) return;
    const int k = (int)(i % A);
    const size_t p = i / A;
    const int y = (int)(p / W), x = (int)(p % W);
    const int yy = y + k / e - r, xx = x + k % e - r;
    float v = 0.0f;
    if (yy >= 0 && yy < H && xx >= 0 && xx < W)
        v = prox[k] * expf(__fdiv_rn(-lab_dist(lab + 3 * p, lab + 3 * ((size_t)yy * W + xx)), gc));
    w[i] = v;
}

// seq/AdaptiveWeightAggregation.cpp:73-111 from precomputed weights:
// one thread per (x, y) and a chunk of d, window in row-major order.
__global__ void __launch_bounds__(256) k_aw_weights(const float *__restrict__ wref, const float *__restrict__ wtar,
                                                    const float *__restrict__ vol, float *__restrict__ out, int W,
                                                    int H, int D, int r, int dir) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const int e = 2 * r + 1, A = e * e;
    const size_t pl = plane(W, H), p = (size_t)y * W + x;
    const int d0 = blockIdx.z * 16, dend = min(D, d0 + 16);
    for (int d = d0; d < dend; d++) {
        const size_t q = (size_t)y * W + wrap1(dir ? x + d : x - d, W);
        float sw = 0.0f, sc = 0.0f;
        int k = 0;
        for (int dy = -r; dy <= r; dy++) {
            const size_t rr = (size_t)d * pl + (size_t)wrap1(y + dy, H) * W;
            for (int dx = -r; dx <= r; dx++, k++) {
                const float wgt = wref[(size_t)A * p + k] * wtar[(size_t)A * q + k];
                sw += wgt;
                sc += vol[rr + wrap1(x + dx, W)] * wgt;
            }
        }
        out[(size_t)d * pl + p] = __fdiv_rn(sc, sw);
    }
}

// ---------------------------------------------------------------- cross arms

// seq/CrossBasedAggregation.cpp:12-100: an arm grows while the previous
// step's max channel delta is below tao, it is shorter than L and the next
// pixel is inside the image; the step that reaches tao is included.  A zero
// arm with an in-image first neighbour becomes 1 (only possible for L < 1).
__device__ __forceinline__ int arm_len(const uint8_t *img, int W, int H, int y, int x, int sy, int sx, int L, int tao,
                                       uint32_t c) {
    int n = 0, delta = 0;
    while (delta < tao && n < L) {
        const int yy = y + sy * (n + 1), xx = x + sx * (n + 1);
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) break;
        n++;
        delta = max_chan_delta(c, pack_rgb(img + 3 * ((size_t)yy * W + xx)));
    }
    if (n == 0) {
        const int yy = y + sy, xx = x + sx;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) n = 1;
    }
    return n;
}

__global__ void __launch_bounds__(256) k_find_cross(const uint8_t *__restrict__ img, uint32_t *__restrict__ arms, int W,
                                                    int H, int L, int tao) {
    const int x = blockIdx.x * 64 + threadIdx.x, y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const uint32_t c = pack_rgb(img + 3 * ((size_t)y * W + x));
    const uint32_t l = arm_len(img, W, H, y, x, 0, -1, L, tao, c);
    const uint32_t r = arm_len(img, W, H, y, x, 0, 1, L, tao, c);
    const uint32_t t = arm_len(img, W, H, y, x, -1, 0, L, tao, c);
    const uint32_t b = arm_len(img, W, H, y, x, 1, 0, L, tao, c);
    arms[(size_t)y * W + x] = l | (r << 8) | (t << 16) | (b << 24);
}

// The same for arms up to 17 with the block's neighbourhood (4 + 2*17 rows
// by 64 + 2*17 columns of packed pixels) staged in LDS.
constexpr int FC_L = 17;
__global__ void __launch_bounds__(256) k_find_cross_tile(const uint8_t *__restrict__ img, uint32_t *__restrict__ arms,
                                                         int W, int H, int L, int tao) {
    __shared__ uint32_t tile[4 + 2 * FC_L][64 + 2 * FC_L + 1];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    constexpr int TH = 4 + 2 * FC_L, TW = 64 + 2 * FC_L;
    for (int i = threadIdx.y * 64 + threadIdx.x; i < TH * TW; i += 256) {
        const int r = i / TW, c = i - r * TW;
        const int yy = y0 - FC_L + r, xx = x0 - FC_L + c;
        tile[r][c] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? pack_rgb(img + 3 * ((size_t)yy * W + xx)) : 0u;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= W || y >= H) return;
    const int tx = threadIdx.x + FC_L, ty = threadIdx.y + FC_L;
    const uint32_t c = tile[ty][tx];
    auto len = [&](int sy, int sx) {
        int n = 0, delta = 0;
        while (delta < tao && n < L) {
            const int yy = y + sy * (n + 1), xx = x + sx * (n + 1);
            if (yy < 0 || yy >= H || xx < 0 || xx >= W) break;
            n++;
            delta = max_chan_delta(c, tile[ty + sy * n][tx + sx * n]);
        }
        if (n == 0) {
            const int yy = y + sy, xx = x + sx;
            if (yy >= 0 && yy < H && xx >= 0 && xx < W) n = 1;
        }
        return (uint32_t)n;
    };
    const uint32_t l = len(0, -1), r = len(0, 1), t = len(-1, 0), b = len(1, 0);
    arms[(size_t)y * W + x] = l | (r << 8) | (t << 16) | (b << 24);
}

// The same with 16 output rows per block (4 per thread), so that the
// 34-row halo is staged once per 16 rows; PACKED reads the packed colour
// image (one load per staged pixel), otherwise the interleaved bytes.
constexpr int FC_TY = 16;
template <bool PACKED>
__global__ void __launch_bounds__(256) k_find_cross_t(const void *__restrict__ src, uint32_t *__restrict__ arms,
                                                      int W, int H, int L, int tao) {
    constexpr int TH = FC_TY + 2 * FC_L, TW = 64 + 2 * FC_L;
    __shared__ uint32_t tile[TH][TW + 1];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * FC_TY;
    for (int i = threadIdx.y * 64 + threadIdx.x; i < TH * TW; i += 256) {
        const int r = i / TW, c = i - r * TW;
        const int yy = y0 - FC_L + r, xx = x0 - FC_L + c;
        uint32_t v = 0;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
            const size_t q = (size_t)yy * W + xx;
            v = PACKED ? static_cast<const uint32_t *>(src)[q] : pack_rgb(static_cast<const uint8_t *>(src) + 3 * q);
        }
        tile[r][c] = v;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x;
    if (x >= W) return;
    const int tx = threadIdx.x + FC_L;
    for (int ly = threadIdx.y; ly < FC_TY; ly += 4) {
        const int y = y0 + ly;
        if (y >= H) break;
        const int ty = ly + FC_L;
        const uint32_t c = tile[ty][tx];
        // arm length: steps while the previous step's delta is below tao and
        // the arm is shorter than L; the step reaching tao is included; the
        // walk stops at the image edge, dist pixels away; a zero arm with an
        // in-image first neighbour becomes 1
        auto len = [&](int sy, int sx, int dist) {
            const int lim = min(L, dist);
            int n = 0, delta = 0;
            while (delta < tao && n < lim) {
                n++;
                delta = max_chan_delta(c, tile[ty + sy * n][tx + sx * n]);
            }
            if (n == 0 && dist > 0) n = 1;
            return (uint32_t)n;
        };
        const uint32_t l = len(0, -1, x), r = len(0, 1, W - 1 - x);
        const uint32_t t = len(-1, 0, y), b = len(1, 0, H - 1 - y);
        arms[(size_t)y * W + x] = l | (r << 8) | (t << 16) | (b << 24);
    }
}

__global__ void k_arms_unpack(const uint32_t *__restrict__ arms, int *__restrict__ horz, int *__restrict__ vert, int P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const uint32_t a = arms[p];
    horz[2 * p] = arm_l(a);
    horz[2 * p + 1] = arm_r(a);
    vert[2 * p] = arm_t(a);
    vert[2 * p + 1] = arm_b(a);
}

__global__ void k_arms_pack(const int *__restrict__ horz, const int *__restrict__ vert, uint32_t *__restrict__ arms,
                            int P) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    arms[p] = (uint32_t)horz[2 * p] | ((uint32_t)horz[2 * p + 1] << 8) | ((uint32_t)vert[2 * p] << 16) |
              ((uint32_t)vert[2 * p + 1] << 24);
}

// ---------------------------------------------------------------- row prefix

// seq/CrossBasedAggregation.cpp:117-125: P[x] = P[x-1] + c[x], serially in
// float along every (d, y) row.  The rows of a [D][H][W] volume are the rows
// of a [D*H][W] matrix; one 64-lane wavefront owns 64 rows, one lane per row,
// so each lane runs its row's additions in the reference's order.  Column
// blocks of 64 are staged through LDS so global loads and stores stay
// row-contiguous (16 lanes x float4 per row segment).
constexpr int PF_ROWS = 64;

__global__ void __launch_bounds__(64) k_hprefix(float *__restrict__ vol, long rows, int W, int vec) {
    __shared__ float tile[PF_ROWS][64 + 1];
    const long r0 = (long)blockIdx.x * PF_ROWS;
    const int lane = threadIdx.x;
    const int nr = (int)min((long)PF_ROWS, rows - r0);
    float acc = 0.0f;
    for (int c0 = 0; c0 < W; c0 += 64) {
        const int nc = min(64, W - c0);
        if (vec && nc == 64) {
            // 16 independent float4 loads per lane, then the LDS stores.
            const int sub = lane >> 4, q = (lane & 15) * 4;
            float4 v[16];
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int i = sub + 4 * k;
                v[k] = i < nr ? *reinterpret_cast<const float4 *>(vol + (r0 + i) * (long)W + c0 + q)
                              : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int k = 0; k < 16; k++) {
                const int i = sub + 4 * k;
                tile[i][q] = v[k].x;
                tile[i][q + 1] = v[k].y;
                tile[i][q + 2] = v[k].z;
                tile[i][q + 3] = v[k].w;
            }
        } else {
            float v[PF_ROWS];
#pragma unroll
            for (int i = 0; i < PF_ROWS; i++) v[i] = (i < nr && lane < nc) ? vol[(r0 + i) * (long)W + c0 + lane] : 0.0f;
#pragma unroll
            for (int i = 0; i < PF_ROWS; i++) tile[i][lane] = v[i];
        }
        __syncthreads();
        if (lane < nr) {
            int j = 0;
            if (c0 == 0) {
                acc = tile[lane][0];
                j = 1;
            }
            for (; j < nc; j++) {
                acc = acc + tile[lane][j];
                tile[lane][j] = acc;
            }
        }
        __syncthreads();
        if (vec && nc == 64) {
            const int sub = lane >> 4, q = (lane & 15) * 4;
            for (int i = sub; i < nr; i += 4) {
                float4 v;
                v.x = tile[i][q];
                v.y = tile[i][q + 1];
                v.z = tile[i][q + 2];
                v.w = tile[i][q + 3];
                *reinterpret_cast<float4 *>(vol + (r0 + i) * (long)W + c0 + q) = v;
            }
        } else {
            for (int i = 0; i < nr; i++)
                if (lane < nc) vol[(r0 + i) * (long)W + c0 + lane] = tile[i][lane];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- vertical sweep

// seq/CrossBasedAggregation.cpp:127-169.  One block covers one d plane tile
// of CV_TX columns x TY output rows and loads everything it needs up front
// with independent, coalesced loads:
//   - the input rows y0-17 .. y0+TY+16, columns x0-18 .. x0+CV_TX+16
//     (clipped to the image) into LDS;
//   - the ref arms at (row, x) and the tar arms at (row, x-d), combined into
//     one packed word of minima per (row, column).
// Segment sums S(row, x) over columns x-l .. x+r (with column 0 dropped when
// x-l-1 < 0, seq/CrossBasedAggregation.cpp:151-156) then come from LDS, and
// each output adds its rows' segments and lengths in the reference's row
// order in float, then divides by the count.
//   PREFIX input (the reference's dataflow): S = P[x+r] - P[max(x-l-1, 0)]
//     of the row prefix P, bit-identical to seq/ cross_stereo_aggregation.
//   DIRECT input (raw costs): S = sum of the costs in the segment, added in
//     column order; no prefix volume, and each float sum has at most 35 terms.
// Arms longer than the 17-row/column halo (possible only with L > 17) are
// handled from global memory.
constexpr int CV_TX = 64;
constexpr int CV_LMAX = 17;
constexpr int CV_COLS = CV_TX + 2 * CV_LMAX + 1;   // x0-18 .. x0+80

template <bool DIRECT, int TY>
__global__ void __launch_bounds__(256) k_cross_tile(const float *__restrict__ in, float *__restrict__ out,
                                                    const uint32_t *__restrict__ aref,
                                                    const uint32_t *__restrict__ atar, int W, int H) {
    constexpr int R = TY + 2 * CV_LMAX;
    constexpr int PER = (R + 3) / 4;                 // rows per thread in the segment pass
    __shared__ float tile[R][CV_COLS + 1];           // input rows; later the segment sums
    __shared__ uint32_t arm[R][CV_TX];               // packed minima: l, r, t, b
    const int x0 = blockIdx.x * CV_TX, y0 = blockIdx.y * TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const int rlo = max(0, y0 - CV_LMAX), rhi = min(H, y0 + TY + CV_LMAX);
    const int nrows = rhi - rlo;
    const int clo = max(0, x0 - CV_LMAX - 1), chi = min(W, x0 + CV_TX + CV_LMAX);
    const int ncols = chi - clo;
    const int tid = threadIdx.x;
    // Phase 1: bulk loads.  Every load of the thread is issued before any
    // LDS store (the counts are compile-time), so they are all in flight
    // together instead of one round trip per element.
    constexpr int NT = (R * CV_COLS + 255) / 256, NA = (R * CV_TX + 255) / 256;
    float tv[NT];
    uint32_t av[NA], bv[NA];
#pragma unroll
    for (int k = 0; k < NT; k++) {
        const int e = tid + 256 * k;
        const int rr = e / CV_COLS, cc = e - rr * CV_COLS;
        tv[k] = (rr < nrows && cc < ncols) ? in[base + (size_t)(rlo + rr) * W + clo + cc] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int e = tid + 256 * k;
        const int rr = min(e >> 6, nrows - 1), cc = e & 63;
        const int x = min(x0 + cc, W - 1);
        const size_t ro = (size_t)(rlo + rr) * W;
        av[k] = aref[ro + x];
        bv[k] = atar[ro + wrap1(x - d, W)];
    }
#pragma unroll
    for (int k = 0; k < NT; k++) {
        const int e = tid + 256 * k;
        const int rr = e / CV_COLS, cc = e - rr * CV_COLS;
        if (rr < R) tile[rr][cc] = tv[k];
    }
#pragma unroll
    for (int k = 0; k < NA; k++) {
        const int e = tid + 256 * k;
        const int rr = e >> 6, cc = e & 63;
        const uint32_t a = av[k], b = bv[k];
        if (rr < R)
            arm[rr][cc] = (uint32_t)min(arm_l(a), arm_l(b)) | ((uint32_t)min(arm_r(a), arm_r(b)) << 8) |
                          ((uint32_t)min(arm_t(a), arm_t(b)) << 16) | ((uint32_t)min(arm_b(a), arm_b(b)) << 24);
    }
    __syncthreads();
    // Phase 2: segment sums of (row, column) into registers.
    const int lx = tid & 63, lg = tid >> 6;
    const int x = x0 + lx;
    const bool colok = x < W;
    const int tx = colok ? wrap1(x - d, W) : 0;
    float sv[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lg + 4 * k;
        float sum = 0.0f;
        if (colok && i < nrows) {
            const uint32_t a = arm[i][lx];
            const int l = arm_l(a), r = arm_r(a);
            const int xl = max(x - l - 1, 0), xr = min(x + r, W - 1);
            if (xl >= clo && xr < chi) {
                if (DIRECT) {
                    for (int c = xl + 1 - clo; c <= xr - clo; c++) sum += tile[i][c];
                } else {
                    sum = tile[i][xr - clo] - tile[i][xl - clo];
                }
            } else {
                const float *rowp = in + base + (size_t)(rlo + i) * W;
                if (DIRECT) {
                    for (int c = xl + 1; c <= xr; c++) sum += rowp[c];
                } else {
                    sum = rowp[xr] - rowp[xl];
                }
            }
        }
        sv[k] = sum;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int i = lg + 4 * k;
        if (i < nrows) tile[i][lx] = sv[k];
    }
    __syncthreads();
    if (!colok) return;
    // Phase 3: each output adds its rows in order.
    for (int ly = lg; ly < TY; ly += 4) {
        const int y = y0 + ly;
        if (y >= H) break;
        const uint32_t a = arm[y - rlo][lx];
        const int t = arm_t(a), b = arm_b(a);
        const int ia = y - t - rlo, ib = y + b - rlo;
        float sum = 0.0f;
        int cnt = 0;
        if (ia >= 0 && ib < nrows) {
            for (int i = ia; i <= ib; i++) {
                const uint32_t ai = arm[i][lx];
                sum += tile[i][lx];
                cnt += arm_l(ai) + arm_r(ai) + 1;
            }
        } else {
            // Vertical arms beyond the halo: rows from global memory.
            for (int row = max(y - t, 0); row <= min(y + b, H - 1); row++) {
                const uint32_t rr = aref[(size_t)row * W + x], rt = atar[(size_t)row * W + tx];
                const int l = min(arm_l(rr), arm_l(rt)), r = min(arm_r(rr), arm_r(rt));
                const int xl = max(x - l - 1, 0), xr = min(x + r, W - 1);
                const float *rowp = in + base + (size_t)row * W;
                float seg = 0.0f;
                if (DIRECT) {
                    for (int c = xl + 1; c <= xr; c++) seg += rowp[c];
                } else {
                    seg = rowp[xr] - rowp[xl];
                }
                sum += seg;
                cnt += l + r + 1;
            }
        }
        out[base + (size_t)y * W + x] = __fdiv_rn(sum, (float)cnt);
    }
}

// The prefix-input sweep streamed down a column strip.  A block owns 64
// columns and a segment of SEG output rows of one d plane and walks the rows
// of [seg-17, seg+SEG+17) in steps of 16: each step's 16 rows are loaded (the
// 99 prefix values a row's segments can touch, held as two registers per
// lane, plus the ref and tar arms of the lane's column), their segment sums
// S and packed (length, top, bottom) words go into a 64-row LDS ring, and
// every output row whose vertical support is complete is emitted: its
// segments added in the reference's row order, then divided by the count
// (seq/CrossBasedAggregation.cpp:127-169, bit-identical).  Each segment sum
// is computed once per row (no tile halo), and the next step's loads are in
// flight while the current outputs are summed.
constexpr int SW_CH = 16;        // rows per step (4 per wave)
constexpr int SW_RING = 64;      // ring rows (power of two, >= SW_CH + 2*17 + SW_CH)
constexpr int SW_SEG = 136;      // output rows per block

// Column j (0..98) of a row segment held by a wave as two registers per
// lane (lanes 0..63: columns 0..63, lanes 0..34: columns 64..98), fetched
// from the owning lane with ds_bpermute.
__device__ __forceinline__ float lane_fetch(float v0, float v1, int j) {
    const int src = (j & 63) << 2;
    const float a = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(v0)));
    const float b = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(v1)));
    return j < 64 ? a : b;
}

__device__ __forceinline__ uint32_t min_u8x4(uint32_t a, uint32_t b) {
    // per-byte minimum of two packed arm words via two 16-bit pair minima
    typedef unsigned short us2 __attribute__((ext_vector_type(2)));
    const uint32_t alo = a & 0x00FF00FFu, blo = b & 0x00FF00FFu;
    const uint32_t ahi = (a >> 8) & 0x00FF00FFu, bhi = (b >> 8) & 0x00FF00FFu;
    us2 x0 = __builtin_bit_cast(us2, alo), y0 = __builtin_bit_cast(us2, blo);
    us2 x1 = __builtin_bit_cast(us2, ahi), y1 = __builtin_bit_cast(us2, bhi);
    const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(x0, y0));
    const uint32_t hi = __builtin_bit_cast(uint32_t, __builtin_elementwise_min(x1, y1));
    return lo | (hi << 8);
}

// This form serves arms longer than CV_LMAX (cross_L > 17): a segment or a
// vertical support longer than the staged window falls back to reading the
// prefix volume and the arm maps directly.
__global__ void __launch_bounds__(256) k_cross_strip(const float *__restrict__ P, float *__restrict__ out,
                                                     const uint32_t *__restrict__ aref,
                                                     const uint32_t *__restrict__ atar, int W, int H, int nx,
                                                     int nseg, int D) {
    __shared__ float2 ring[SW_RING][CV_TX];   // x: segment sum, y: bits of (len | top<<10 | bottom<<18)
    // Work order (segment, d, x-tile) with x fastest, cut into one
    // contiguous range per XCD: x-neighbours (which share 35 prefix columns)
    // run on the same L2.
    const int wk = xcd_order(blockIdx.x, nx * nseg * D);
    const int xt = wk % nx, rest = wk / nx;
    const int d = rest % D, sg = rest / D;
    const int x0 = xt * CV_TX, s0 = sg * SW_SEG;
    const int s1 = min(H, s0 + SW_SEG);
    const size_t base = (size_t)d * plane(W, H);
    const int rlo = max(0, s0 - CV_LMAX), rhi = min(H, s1 + CV_LMAX);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // wave-uniform: scalar row bases
    const int x = x0 + lane;
    const bool colok = x < W;
    const int xc = min(x, W - 1);
    const int tx = wrap1(xc - d, W);
    const int cs = x0 - CV_LMAX - 1;
    const int c0 = cs + lane, c1 = cs + 64 + lane;
    const bool ok0 = c0 >= 0 && c0 < W, ok1 = lane < CV_COLS - 64 && c1 >= 0 && c1 < W;
    const float *Pb = P + base;
    float p0[4], p1[4];
    uint32_t ar[4], at[4];
    auto load_rows = [&](int r) {
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int row = min(r + w + 4 * k, rhi - 1);
            const size_t ro = (size_t)row * W;
            p0[k] = ok0 ? Pb[ro + c0] : 0.0f;
            p1[k] = ok1 ? Pb[ro + c1] : 0.0f;
            ar[k] = aref[ro + xc];
            at[k] = atar[ro + tx];
        }
    };
    int out_next = s0;   // first output row not yet written
    load_rows(rlo);
    for (int r = rlo; r < rhi; r += SW_CH) {
        // segment sums of rows r .. r+15 into the ring
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int row = r + w + 4 * k;
            const uint32_t a = min_u8x4(ar[k], at[k]);
            const int l = arm_l(a), rr = arm_r(a);
            float sv;
            if (l <= CV_LMAX && rr <= CV_LMAX) {
                const int xl = max(xc - l - 1, 0), xr = min(xc + rr, W - 1);
                sv = lane_fetch(p0[k], p1[k], xr - cs) - lane_fetch(p0[k], p1[k], xl - cs);
            } else {
                // horizontal arm longer than the segment window (L > 17)
                const float *Pr = Pb + (size_t)min(row, rhi - 1) * W;
                sv = Pr[min(xc + rr, W - 1)] - Pr[max(xc - l - 1, 0)];
            }
            if (row < rhi) {
                const float2 v = make_float2(
                    sv, __uint_as_float((uint32_t)(l + rr + 1) | ((a >> 16) & 0xFFu) << 10 | (a >> 24) << 18));
                ring[row & (SW_RING - 1)][lane] = v;
            }
        }
        __syncthreads();
        const int rnext = r + SW_CH;
        if (rnext < rhi) load_rows(rnext);
        // outputs whose rows are all in the ring
        const int avail = min(rnext, rhi);
        const int out_end = avail >= rhi ? s1 : min(s1, avail - CV_LMAX);
        if (colok)
            for (int y = out_next + w; y < out_end; y += 4) {
                const uint32_t word = __float_as_uint(ring[y & (SW_RING - 1)][lane].y);
                const int t = (word >> 10) & 0xFF, b = word >> 18;
                float sum = 0.0f;
                int cnt = 0;
                if (t <= CV_LMAX && b <= CV_LMAX) {
                    // rows y-t .. y+b, top to bottom (the reference's order),
                    // four reads issued together
                    const int n = t + b + 1, y0r = y - t;
                    int i = 0;
                    for (; i + 4 <= n; i += 4) {
                        const float2 e0 = ring[(y0r + i) & (SW_RING - 1)][lane];
                        const float2 e1 = ring[(y0r + i + 1) & (SW_RING - 1)][lane];
                        const float2 e2 = ring[(y0r + i + 2) & (SW_RING - 1)][lane];
                        const float2 e3 = ring[(y0r + i + 3) & (SW_RING - 1)][lane];
                        sum += e0.x;
                        sum += e1.x;
                        sum += e2.x;
                        sum += e3.x;
                        cnt += (int)((__float_as_uint(e0.y) & 0x3FF) + (__float_as_uint(e1.y) & 0x3FF) +
                                     (__float_as_uint(e2.y) & 0x3FF) + (__float_as_uint(e3.y) & 0x3FF));
                    }
                    for (; i < n; i++) {
                        const float2 e = ring[(y0r + i) & (SW_RING - 1)][lane];
                        sum += e.x;
                        cnt += __float_as_uint(e.y) & 0x3FF;
                    }
                } else {
                    // vertical arm longer than the ring window (L > 17)
                    for (int row = max(y - t, 0); row <= min(y + b, H - 1); row++) {
                        const uint32_t rr2 = aref[(size_t)row * W + x], rt = atar[(size_t)row * W + tx];
                        const int l2 = min(arm_l(rr2), arm_l(rt)), r2 = min(arm_r(rr2), arm_r(rt));
                        const float *Pr = Pb + (size_t)row * W;
                        sum += Pr[min(x + r2, W - 1)] - Pr[max(x - l2 - 1, 0)];
                        cnt += l2 + r2 + 1;
                    }
                }
                out[base + (size_t)y * W + x] = __fdiv_rn(sum, (float)cnt);
            }
        out_next = max(out_next, out_end);
        __syncthreads();
    }
}

// Element at a 32-bit byte offset from a uniform base, so that the load can
// take the scalar-base form with a 32-bit per-lane offset.
template <typename T>
__device__ __forceinline__ T at_byte(const T *base, uint32_t off) {
    return *(const T *)((const char *)base + off);
}


// The sweep for arms of at most CV_LMAX (seq/CrossBasedAggregation.cpp:127-169,
// bit-identical).  A block of NW waves owns 64 columns and SEG output rows of
// one d plane and walks rows [s0-17, s0+SEG+17) in steps of CH; every wave
// forms CH/NW rows of segment sums per step and then computes the output rows
// whose support is complete, NW rows apart.
//  - Segment sums: a row's 99 prefix values (columns x0-18 .. x0+80) are
//    loaded coalesced, two registers per lane, one step ahead; each lane
//    picks its segment's two ends with ds_bpermute, so every segment sum is
//    formed once (no tile halo).  Loads are buffer loads: the plane and map
//    bases sit in scalar registers, the row as a scalar offset, the column as
//    a fixed per-lane offset.
//  - The ring keeps R = CH + 34 live rows and repeats its first 34 rows past
//    its end, so the rows y-17 .. y+17 of any output sit at consecutive
//    positions and every read of a vertical sum is an immediate offset from
//    one per-lane base (two rows per ds_read2st64_b64).
//  - A ring entry is (segment sum, len | t << 16 | b << 24): the low 16 bits
//    of the plain 32-bit sum of the entries over a support are its count (at
//    most 35 rows of at most 35 columns), so no field is unpacked per row.
//  - A support of n rows is summed as n % 4 single rows and then n / 4
//    groups of four, in row order from +0.0 (the reference's order); the
//    quotient uses div_count.
//  - A step's outputs (at most CH/NW rows per wave) stay in registers and are
//    stored (non-temporal) after the next step's segment sums, before that
//    step's loads: loads and stores share one completion counter, so a wait
//    for a step's loads then never waits for the stores of the step before.
//    The rows a step past the image bottom adds (up to 17) are computed
//    after the walk.
// Eight waves share one ring, which keeps six waves per SIMD resident at
// 51 KB of LDS per block.
template <int CH, int NW>
__global__ void __launch_bounds__(64 * NW) k_cross_sweep(const float *__restrict__ P, float *__restrict__ out,
                                                         const uint32_t *__restrict__ aref,
                                                         const uint32_t *__restrict__ atar, int W, int H, int nx,
                                                         int nseg, int seg, int D, int xg) {
    constexpr int R = CH + 2 * CV_LMAX;        // rows live at once
    constexpr int RR = R + 2 * CV_LMAX;        // stored rows: the first 34 repeated past the end
    constexpr int KR = CH / NW;                // rows per wave per step
    static_assert(KR * NW == CH, "step rows must split evenly over the waves");
    __shared__ float2 ring[RR][CV_TX];
    // Work order: row segment, then groups of xg column tiles, then d, then
    // the tile within its group; cut into one contiguous range per XCD, so an
    // XCD's blocks share the arm maps of a few tiles across many d (and the
    // prefix columns of neighbouring tiles of one d) in its L2.
    const int wk = xcd_order(blockIdx.x, nx * nseg * D);
    const int per_seg = nx * D;
    const int sg = wk / per_seg, wi = wk - sg * per_seg;
    const int g = wi / (xg * D), wj = wi - g * (xg * D);
    const int gsz = min(xg, nx - g * xg);
    const int d = wj / gsz, xt = g * xg + (wj - d * gsz);
    const int x0 = xt * CV_TX, s0 = sg * seg;
    const int s1 = min(H, s0 + seg);
    const size_t base = (size_t)d * plane(W, H);
    const int rlo = max(0, s0 - CV_LMAX), rhi = min(H, s1 + CV_LMAX);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = x0 + lane;
    const bool colok = x < W;
    const int xc = min(x, W - 1);
    const int tx = wrap1(xc - d, W);
    const int cs = x0 - CV_LMAX - 1;
    // The 99 prefix values a row's segments can touch (columns cs .. cs+98)
    // are held as two registers per lane: lo = column cs+j on lane j, and
    // hi = column cs+64+j on lanes j < 47 (cs+64 .. cs+98 on j <= 34), column
    // cs+j on lanes j >= 47.  A left end cs+jl with jl in [lane, lane+17]
    // then lies in lo for lanes below 47 and in hi (source lane jl mod 64)
    // for the others; a right end jr in [lane+18, lane+35] lies in lo for
    // lanes below 29 and in hi for the others, so the register is a fixed
    // choice per lane.  hi on lanes 35 .. 46 is never read: those lanes load
    // their lo column again, which keeps every load inside the window.
    // Segment ends are always columns inside the image, so a lane whose
    // column lies outside it loads any in-image column instead.
    const int c0 = cs + lane, c1 = cs + lane + (lane < CV_COLS - 64 ? 64 : 0);
    const uint32_t oc0 = (uint32_t)min(max(c0, 0), W - 1) * 4u, oc1 = (uint32_t)min(max(c1, 0), W - 1) * 4u;
    // blocks whose 99 columns all lie inside the image need no clamping
    const bool interior = cs >= 0 && cs + CV_COLS <= W;
    const uint32_t oxc = (uint32_t)xc * 4u, otx = (uint32_t)tx * 4u;
    const bool hi_l = lane >= 47, hi_r = lane >= 29;
    // one plane is at most 2^32 bytes (W*H <= 2^30), so every offset fits 32 bits
    const uint64_t pb64 = (uint64_t)W * (uint64_t)H * 4u;
    const int pbytes = (int)(pb64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)pb64);
    const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void *)(P + base), 0, pbytes, BUF_DWORD3);
    const auto rsR = __builtin_amdgcn_make_buffer_rsrc((void *)aref, 0, pbytes, BUF_DWORD3);
    const auto rsT = __builtin_amdgcn_make_buffer_rsrc((void *)atar, 0, pbytes, BUF_DWORD3);
    // A step's loads are issued one step before its segment sums are formed.
    float p0[KR], p1[KR];
    uint32_t ar[KR], at[KR];
    auto load_rows = [&](int r) {
#pragma unroll
        for (int k = 0; k < KR; k++) {
            const uint32_t ro = (uint32_t)min(r + w + NW * k, rhi - 1) * (uint32_t)W * 4u;
            p0[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsP, oc0, ro, 0));
            p1[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsP, oc1, ro, 0));
            ar[k] = __builtin_amdgcn_raw_buffer_load_b32(rsR, oxc, ro, 0);
            at[k] = __builtin_amdgcn_raw_buffer_load_b32(rsT, otx, ro, 0);
        }
    };
    int out_next = s0;
    int pr = 0;   // ring position of row r: (r - rlo) mod R
    load_rows(rlo);
    // segment sums of one step's rows into the ring; INTERIOR: every column
    // of the block's window lies in the image (no clamps, fixed registers)
    auto produce = [&](auto interior_tag, int r) {
        constexpr bool INTERIOR = decltype(interior_tag)::value;
#pragma unroll
        for (int k = 0; k < KR; k++) {
            const uint32_t a = ar[k], c = at[k];
            const uint32_t l = min(a & 0xFFu, c & 0xFFu), rr = min((a >> 8) & 0xFFu, (c >> 8) & 0xFFu);
            const uint32_t t = min((a >> 16) & 0xFFu, (c >> 16) & 0xFFu), b = min(a >> 24, c >> 24);
            // byte addresses of the segment's ends relative to cs (ds_bpermute
            // takes the source lane from address bits 7:2, so 4j selects lane
            // j % 64)
            const int jl4 = INTERIOR ? 4 * (lane + CV_LMAX) - 4 * (int)l : 4 * (max(xc - (int)l - 1, 0) - cs);
            const int jr4 = INTERIOR ? 4 * (int)rr + 4 * (lane + CV_LMAX + 1) : 4 * (min(xc + (int)rr, W - 1) - cs);
            const int bl = __builtin_amdgcn_ds_bpermute(jl4, __float_as_int(p0[k]));
            const int bh = __builtin_amdgcn_ds_bpermute(jl4, __float_as_int(p1[k]));
            const int el = __builtin_amdgcn_ds_bpermute(jr4, __float_as_int(p0[k]));
            const int eh = __builtin_amdgcn_ds_bpermute(jr4, __float_as_int(p1[k]));
            // hi holds column cs+64+j for j < 47: a clamped end below 64
            // reads lo, one at or above 64 reads hi
            const bool uh_l = INTERIOR ? hi_l : jl4 >= 256, uh_r = INTERIOR ? hi_r : jr4 >= 256;
            const float sv = __int_as_float(uh_r ? eh : el) - __int_as_float(uh_l ? bh : bl);
            if (r + w + NW * k < rhi) {
                const float2 v = make_float2(sv, __uint_as_float((l + rr + 1) | t << 16 | b << 24));
                int p = pr + w + NW * k;
                if (p >= R) p -= R;
                ring[p][lane] = v;
                if (p < 2 * CV_LMAX) ring[p + R][lane] = v;
            }
        }
    };
    // The output whose row y-17 sits at ring position bm; wd is the entry
    // word of row y itself.
    auto out_row = [&](int bm, uint32_t wd) -> float {
        const int t = (wd >> 16) & 0xFF, b = wd >> 24;
        const int n = t + b + 1, lead = n & 3, q = n >> 2;
        const float2 *s = &ring[bm][lane] + (CV_LMAX - t) * CV_TX;
        float sum = 0.0f;
        uint32_t cw = 0;
        if (lead) {
            // the lead rows in one round trip; a row past the lead adds +0.0,
            // which leaves the (never -0.0) sum unchanged
            const float2 a0 = s[0], a1 = s[CV_TX], a2 = s[2 * CV_TX];
            const bool u1 = lead > 1, u2 = lead > 2;
            sum = ((sum + a0.x) + (u1 ? a1.x : 0.0f)) + (u2 ? a2.x : 0.0f);
            cw = __float_as_uint(a0.y) + (u1 ? __float_as_uint(a1.y) : 0u) + (u2 ? __float_as_uint(a2.y) : 0u);
        }
        s += lead * CV_TX;
        if (q >= 1) {
            const float2 e0 = s[0], e1 = s[CV_TX], e2 = s[2 * CV_TX], e3 = s[3 * CV_TX];
            sum = (((sum + e0.x) + e1.x) + e2.x) + e3.x;
            cw += __float_as_uint(e0.y) + __float_as_uint(e1.y) + __float_as_uint(e2.y) + __float_as_uint(e3.y);
        }
        // supports of 8 rows or more (rare): further groups of four
#pragma unroll 1
        for (int gi = 1; gi < q; gi++) {
            const float2 *gp = s + 4 * gi * CV_TX;
            const float2 e0 = gp[0], e1 = gp[CV_TX], e2 = gp[2 * CV_TX], e3 = gp[3 * CV_TX];
            sum = (((sum + e0.x) + e1.x) + e2.x) + e3.x;
            cw += __float_as_uint(e0.y) + __float_as_uint(e1.y) + __float_as_uint(e2.y) + __float_as_uint(e3.y);
        }
        return div_count(sum, cw & 0xFFFFu);
    };
    // non-temporal stores: the aggregated volume streams out to the next
    // kernel (measured 1.03 -> 1.00 ms here and 0.30 -> 0.28 ms for the WTA
    // that reads it, at 1080p/192)
    // (buffer stores: the row as a scalar offset; cache policy 2 = nt)
    const auto rsO = __builtin_amdgcn_make_buffer_rsrc((void *)(out + base), 0, pbytes, BUF_DWORD3);
    auto store_row = [&](int y, float v) {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), rsO, (uint32_t)x * 4u, (uint32_t)y * (uint32_t)W * 4u,
                                              2);
    };
    float res[KR];                 // this wave's outputs of the last step:
    int res_y = 0, res_end = 0;    // rows res_y + NW*i below res_end
    for (int r = rlo; r < rhi; r += CH) {
        if (interior)
            produce(std::true_type{}, r);
        else
            produce(std::false_type{}, r);
        __syncthreads();
        if (colok) {
#pragma unroll
            for (int i = 0; i < KR; i++)
                if (res_y + NW * i < res_end) store_row(res_y + NW * i, res[i]);
        }
        const int rnext = r + CH;
        if (rnext < rhi) load_rows(rnext);
        const int avail = min(rnext, rhi);
        const int out_end = avail >= rhi ? s1 : min(s1, avail - CV_LMAX);
        // at most CH rows a step; the rest of the last step's rows follow the walk
        const int oe = min(out_end, out_next + CH);
        res_y = out_next + w;
        res_end = oe;
        if (colok) {
            // ring position of row y-17 for the wave's first output row (rows
            // below rlo never enter a support), then advanced with y; each
            // output row's entry word is read one row ahead
            int bm = (out_next + w - CV_LMAX - rlo + R) % R;
            uint32_t word = __float_as_uint(ring[bm + CV_LMAX][lane].y);
#pragma unroll
            for (int i = 0; i < KR; i++) {
                if (res_y + NW * i < oe) {
                    int bn = bm + NW;
                    if (bn >= R) bn -= R;
                    const uint32_t next = __float_as_uint(ring[bn + CV_LMAX][lane].y);
                    res[i] = out_row(bm, word);
                    bm = bn;
                    word = next;
                }
            }
        }
        out_next = max(out_next, oe);
        pr += CH;
        if (pr >= R) pr -= R;
        __syncthreads();
    }
    if (colok) {
#pragma unroll
        for (int i = 0; i < KR; i++)
            if (res_y + NW * i < res_end) store_row(res_y + NW * i, res[i]);
        // rows past the last step's CH outputs (a segment at the image
        // bottom); every row of their supports is still in the ring
        int bm = (out_next + w - CV_LMAX - rlo + R) % R;
        for (int y = out_next + w; y < s1; y += NW) {
            store_row(y, out_row(bm, __float_as_uint(ring[bm + CV_LMAX][lane].y)));
            bm += NW;
            if (bm >= R) bm -= R;
        }
    }
}

// The vertical sweep for arms of at most CV_LMAX fused with the disparity
// selection (seq/CrossBasedAggregation.cpp:127-169, then seq/WTA.cpp:10-39
// for both views), so that the aggregated volume is never stored.  A block
// of FW_NW waves owns 64 columns x FW_SEG rows of the frame and walks every
// d.  Per d it forms the segment sums of the rows its outputs' supports can
// reach (s0-17 .. s0+FW_SEG+16, all live at once in LDS), then each wave sums
// its output rows s0+w+8i exactly as k_cross_sweep does, and compares the
// value instead of storing it:
//  - left view: a running best (value, d) per output in registers, replaced
//    only by a strictly smaller value (d = 0 first: the lowest d wins ties);
//  - right view: the output at column x and disparity d is a candidate for
//    right pixel (x-d) mod W.  Lane i keeps the running best of right pixel
//    x0+i-d; before each next d the state moves up one lane (DPP wave
//    shift), lane 0 starts empty, and the state leaving lane 63, which this
//    block sees no more, is merged into the frame's right-view buffer with a
//    64-bit atomic min of (value bits << 32 | d): non-negative floats order
//    as their bits, and the lowest d wins ties.  After the last d every
//    lane's state is merged.
// The next step's rows are loaded during the current step's sums.  The
// right-view buffer starts at all ones (k_fill_u64).
constexpr int FW_SEG = 64, FW_NW = 8;
constexpr int FW_ROWS = FW_SEG + 2 * CV_LMAX;       // rows live at once
constexpr int FW_RING = FW_ROWS + 8;                // + rows a short support's grouped reads pass
constexpr int FW_KP = (FW_ROWS + FW_NW - 1) / FW_NW;   // rows each wave forms per d
constexpr int FW_KO = FW_SEG / FW_NW;                  // output rows per wave

__device__ __forceinline__ unsigned long long fw_pack(float v, int d) {
    return (unsigned long long)__float_as_uint(v) << 32 | (unsigned)d;
}

template <bool WANT_R>
__global__ void __launch_bounds__(64 * FW_NW) k_cross_sweep_wta(const float *__restrict__ P,
                                                                const uint32_t *__restrict__ aref,
                                                                const uint32_t *__restrict__ atar,
                                                                uint8_t *__restrict__ dl,
                                                                unsigned long long *__restrict__ r2l, int W, int H,
                                                                int D, int nx, int nseg) {
    __shared__ float2 ring[FW_RING][CV_TX];
    // segments fastest: an XCD's blocks hold whole column strips, whose
    // neighbouring segments read each other's halo rows (at about the same
    // d) from its L2
    const int wk = xcd_order(blockIdx.x, nx * nseg);
    const int xt = wk / nseg, sg = wk - xt * nseg;
    const int x0 = xt * CV_TX, s0 = sg * FW_SEG, s1 = min(H, s0 + FW_SEG);
    const int rlo = max(0, s0 - CV_LMAX), rhi = min(H, s1 + CV_LMAX);
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int x = x0 + lane;
    const bool colok = x < W;
    const int xc = min(x, W - 1);
    const int cs = x0 - CV_LMAX - 1;
    // prefix columns cs .. cs+98 in two registers per lane (k_cross_sweep)
    const int c0 = cs + lane, c1 = cs + lane + (lane < CV_COLS - 64 ? 64 : 0);
    const uint32_t oc0 = (uint32_t)min(max(c0, 0), W - 1) * 4u, oc1 = (uint32_t)min(max(c1, 0), W - 1) * 4u;
    const bool interior = cs >= 0 && cs + CV_COLS <= W;
    const uint32_t oxc = (uint32_t)xc * 4u;
    const bool hi_l = lane >= 47, hi_r = lane >= 29;
    const uint64_t pb64 = (uint64_t)W * (uint64_t)H * 4u;
    const int pbytes = (int)(pb64 > 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)pb64);
    const size_t pl = plane(W, H);
    const auto rsR = __builtin_amdgcn_make_buffer_rsrc((void *)aref, 0, pbytes, BUF_DWORD3);
    const auto rsT = __builtin_amdgcn_make_buffer_rsrc((void *)atar, 0, pbytes, BUF_DWORD3);
    auto row_off = [&](int k) { return (uint32_t)min(rlo + w + FW_NW * k, rhi - 1) * (uint32_t)W * 4u; };
    // the reference view's arms do not depend on d
    uint32_t ar[FW_KP];
#pragma unroll
    for (int k = 0; k < FW_KP; k++) ar[k] = __builtin_amdgcn_raw_buffer_load_b32(rsR, oxc, row_off(k), 0);
    float p0[FW_KP], p1[FW_KP];
    uint32_t at[FW_KP];
    auto load_d = [&](int d) {
        const auto rsP = __builtin_amdgcn_make_buffer_rsrc((void *)(P + (size_t)d * pl), 0, pbytes, BUF_DWORD3);
        const uint32_t otx = (uint32_t)wrap1(xc - d, W) * 4u;
#pragma unroll
        for (int k = 0; k < FW_KP; k++) {
            const uint32_t ro = row_off(k);
            p0[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsP, oc0, ro, 0));
            p1[k] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rsP, oc1, ro, 0));
            at[k] = __builtin_amdgcn_raw_buffer_load_b32(rsT, otx, ro, 0);
        }
    };
    // segment sums of this d's rows into the ring (the entry format and the
    // arithmetic of k_cross_sweep)
    auto produce = [&](auto interior_tag) {
        constexpr bool INTERIOR = decltype(interior_tag)::value;
#pragma unroll
        for (int k = 0; k < FW_KP; k++) {
            const uint32_t a = ar[k], c = at[k];
            const uint32_t l = min(a & 0xFFu, c & 0xFFu), rr = min((a >> 8) & 0xFFu, (c >> 8) & 0xFFu);
            const uint32_t t = min((a >> 16) & 0xFFu, (c >> 16) & 0xFFu), b = min(a >> 24, c >> 24);
            const int jl4 = INTERIOR ? 4 * (lane + CV_LMAX) - 4 * (int)l : 4 * (max(xc - (int)l - 1, 0) - cs);
            const int jr4 = INTERIOR ? 4 * (int)rr + 4 * (lane + CV_LMAX + 1) : 4 * (min(xc + (int)rr, W - 1) - cs);
            const int bl = __builtin_amdgcn_ds_bpermute(jl4, __float_as_int(p0[k]));
            const int bh = __builtin_amdgcn_ds_bpermute(jl4, __float_as_int(p1[k]));
            const int el = __builtin_amdgcn_ds_bpermute(jr4, __float_as_int(p0[k]));
            const int eh = __builtin_amdgcn_ds_bpermute(jr4, __float_as_int(p1[k]));
            const bool uh_l = INTERIOR ? hi_l : jl4 >= 256, uh_r = INTERIOR ? hi_r : jr4 >= 256;
            const float sv = __int_as_float(uh_r ? eh : el) - __int_as_float(uh_l ? bh : bl);
            const int row = rlo + w + FW_NW * k;
            if (row < rhi) ring[row - rlo][lane] = make_float2(sv, __uint_as_float((l + rr + 1) | t << 16 | b << 24));
        }
    };
    // the aggregated value of the output whose own row sits at ring position py
    auto out_row = [&](int py) -> float {
        const uint32_t wd = __float_as_uint(ring[py][lane].y);
        const int t = (wd >> 16) & 0xFF, b = wd >> 24;
        const int n = t + b + 1, lead = n & 3, q = n >> 2;
        const float2 *s = &ring[py - t][lane];
        float sum = 0.0f;
        uint32_t cw = 0;
        if (lead) {
            const float2 a0 = s[0], a1 = s[CV_TX], a2 = s[2 * CV_TX];
            const bool u1 = lead > 1, u2 = lead > 2;
            sum = ((sum + a0.x) + (u1 ? a1.x : 0.0f)) + (u2 ? a2.x : 0.0f);
            cw = __float_as_uint(a0.y) + (u1 ? __float_as_uint(a1.y) : 0u) + (u2 ? __float_as_uint(a2.y) : 0u);
        }
        s += lead * CV_TX;
#pragma unroll 1
        for (int gi = 0; gi < q; gi++) {
            const float2 *gp = s + 4 * gi * CV_TX;
            const float2 e0 = gp[0], e1 = gp[CV_TX], e2 = gp[2 * CV_TX], e3 = gp[3 * CV_TX];
            sum = (((sum + e0.x) + e1.x) + e2.x) + e3.x;
            cw += __float_as_uint(e0.y) + __float_as_uint(e1.y) + __float_as_uint(e2.y) + __float_as_uint(e3.y);
        }
        return div_count(sum, cw & 0xFFFFu);
    };
    const float INF = __uint_as_float(0x7F800000u);
    float bL[FW_KO], bR[FW_KO];
    int dL[FW_KO], dR[FW_KO];
#pragma unroll
    for (int i = 0; i < FW_KO; i++) {
        bL[i] = bR[i] = INF;
        dL[i] = dR[i] = 0;
    }
    load_d(0);
    for (int d = 0; d < D; d++) {
        if (interior)
            produce(std::true_type{});
        else
            produce(std::false_type{});
        __syncthreads();
        if (d + 1 < D) load_d(d + 1);
#pragma unroll
        for (int i = 0; i < FW_KO; i++) {
            const int y = s0 + w + FW_NW * i;
            if (y >= s1) break;
            const float v = colok ? out_row(y - rlo) : INF;
            if (d == 0 || v < bL[i]) {
                bL[i] = v;
                dL[i] = d;
            }
            if (WANT_R && v < bR[i]) {
                bR[i] = v;
                dR[i] = d;
            }
        }
        if (WANT_R) {
            // the state of right pixel x0+63-d leaves the block; the others
            // move up one lane, and lane 0 starts right pixel x0-d-1
#pragma unroll
            for (int i = 0; i < FW_KO; i++) {
                const int y = s0 + w + FW_NW * i;
                if (y >= s1) break;
                if (lane == 63 && bR[i] != INF)
                    atomicMin(&r2l[(size_t)y * W + wrapn(x0 + 63 - d, W)], fw_pack(bR[i], dR[i]));
                bR[i] = __int_as_float(__builtin_amdgcn_update_dpp((int)0x7F800000, __float_as_int(bR[i]), 0x138, 0xf,
                                                                   0xf, false));   // wave_shr:1
                dR[i] = __builtin_amdgcn_update_dpp(0, dR[i], 0x138, 0xf, 0xf, false);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < FW_KO; i++) {
        const int y = s0 + w + FW_NW * i;
        if (y >= s1) break;
        if (colok) dl[(size_t)y * W + x] = (uint8_t)dL[i];
        // after the last shift lane i holds right pixel x0+i-D
        if (WANT_R && bR[i] != INF) atomicMin(&r2l[(size_t)y * W + wrapn(x0 + lane - D, W)], fw_pack(bR[i], dR[i]));
    }
}

__global__ void k_fill_u64(unsigned long long *__restrict__ p, unsigned long long v, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// D_R from the merged right-view bests: the low 32 bits hold d.
__global__ void k_r2l_finish(const unsigned long long *__restrict__ r2l, uint8_t *__restrict__ dr, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dr[i] = (uint8_t)(r2l[i] & 0xFFu);
}

}  // namespace rts

using namespace rts;

// Shape of the short-arm sweep: 32 rows per step over 8 waves; a block owns
// a column strip of XS_SEG_MIN.. rows, cut so that the grid holds about 4096
// blocks (whole columns for 1080p at D >= 72).  Work groups of XS_GROUP
// column tiles: an XCD's blocks walk every d of about four tiles, which
// halves the sweep's fetches from memory (the arm maps stay in that XCD's
// L2) at the same speed.
constexpr int XS_CH = 32, XS_NW = 8, XS_SEG_MIN = 64, XS_GROUP = 4;
static int xs_segments(int W, int H, int D) {
    const long strips = (long)div_up(W, CV_TX) * D;
    const long most = H / XS_SEG_MIN > 1 ? H / XS_SEG_MIN : 1;
    const long want = (4096L + strips - 1) / strips;
    return (int)(want < most ? want : most);
}

namespace rts {
int find_cross_packed(const uint32_t *pcol, uint32_t *arms, int W, int H, int L, int tao, hipStream_t s) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size %dx%d", W, H);
    RTS_CHECK(L >= 0 && L <= FC_L, "arm limit %d must be in [0, %d]", L, FC_L);
    k_find_cross_t<true><<<dim3(div_up(W, 64), div_up(H, FC_TY)), dim3(64, 4), 0, s>>>(pcol, arms, W, H, L, tao);
    RTS_LAUNCHED();
    return RTS_OK;
}
}  // namespace rts

template <bool DIRECT>
static int launch_cross_tile(const float *in, float *out, const uint32_t *ar, const uint32_t *at, int W, int H, int D,
                             hipStream_t s) {
    constexpr int ty = 32;
    RTS_CHECK(div_up(H, ty) <= 65535, "height %d too large", H);
    dim3 grid(div_up(W, CV_TX), div_up(H, ty), D);
    k_cross_tile<DIRECT, ty><<<grid, 256, 0, s>>>(in, out, ar, at, W, H);
    RTS_LAUNCHED();
    return RTS_OK;
}

extern "C" {

rts_status rts_cross_fused(const float *cost, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W,
                           int H, int D, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(cost != out, "cross aggregation is out of place");
    return launch_cross_tile<true>(cost, out, arms_ref, arms_tar, W, H, D, S(stream));
}

rts_status rts_box(const float *in, float *out, int W, int H, int D, int r, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r < W && r < H, "box radius %d does not fit the image", r);
    RTS_CHECK(in != out, "rts_box is out of place");
    if (r <= BOX_RT) {
        const dim3 grid(div_up(W, 64), div_up(H, BOXR_TY), D);
        hipStream_t s = S(stream);
        switch (r) {
            case 0: k_box_rw<0><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 1: k_box_rw<1><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 2: k_box_rw<2><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 3: k_box_rw<3><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 4: k_box_rw<4><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 5: k_box_rw<5><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 6: k_box_rw<6><<<grid, 256, 0, s>>>(in, out, W, H); break;
            case 7: k_box_rw<7><<<grid, 256, 0, s>>>(in, out, W, H); break;
            default: k_box_rw<BOX_RT><<<grid, 256, 0, s>>>(in, out, W, H); break;
        }
    } else if (r <= BOX_RMAX) {
        dim3 grid(div_up(W, 64), div_up(H, BOX_TY), D), block(64, 4);
        k_box_f32<<<grid, block, 0, S(stream)>>>(in, out, W, H, r);
    } else {
        dim3 grid(div_up(W, 64), div_up(H, 4), D), block(64, 4);
        k_box_f32_global<<<grid, block, 0, S(stream)>>>(in, out, W, H, r);
    }
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_box_int(const float *in, float *out, int W, int H, int D, int r, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r <= BOXI_RMAX && r < W && r < H, "box radius %d must be in [0, %d] and fit the image", r,
              BOXI_RMAX);
    RTS_CHECK(in != out, "rts_box_int is out of place");
    const int ntx = div_up(W, 64);
    // segments of at least 16 rows, enough of them for about 16384 waves
    // (two rounds of the chip at eight waves per SIMD)
    const long cols = (long)ntx * D;
    const int most = H / 16 > 1 ? H / 16 : 1;
    const int nseg = (int)std::min<long>(most, std::max<long>(1, (16384 + cols - 1) / cols));
    const int seg = div_up(H, nseg);
    const long nwork = cols * nseg;
    RTS_CHECK(nwork < (1L << 31) - 4, "volume too large");
    const size_t lds = sizeof(uint32_t) * 4 * (2 * r + 1) * 64;
    k_box_int<<<(unsigned)div_up(nwork, 4L), 256, lds, S(stream)>>>(in, out, W, H, r, ntx, nseg, seg, (int)nwork);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_prox_weight(float *prox, float gamma_p, int r, void *stream) {
    RTS_CHECK(r >= 0 && r <= 64, "AW radius %d out of range", r);
    const int A = (2 * r + 1) * (2 * r + 1);
    k_prox<<<div_up(A, 256), 256, 0, S(stream)>>>(prox, gamma_p, r);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_compute_weights(float *weights, const float *prox, const uint8_t *lab, float gamma_c, int W, int H,
                               int r, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && r >= 0 && r <= 64, "bad AW weight arguments");
    const size_t n = (size_t)(2 * r + 1) * (2 * r + 1) * W * H;
    RTS_CHECK(n / 256 < (size_t)1 << 31, "weight tensor too large");
    k_weights<<<(unsigned)((n + 255) / 256), 256, 0, S(stream)>>>(weights, prox, lab, gamma_c, W, H, r);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_aw_weights(const float *wref, const float *wtar, const float *vol, float *out, int W, int H, int D,
                          int r, int dir, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r < W && r < H, "AW radius %d does not fit the image", r);
    RTS_CHECK(vol != out, "AW aggregation is out of place");
    dim3 grid(div_up(W, 64), div_up(H, 4), div_up(D, 16)), block(64, 4);
    RTS_CHECK(dir == RTS_L2R || dir == RTS_R2L, "bad direction %d", dir);
    k_aw_weights<<<grid, block, 0, S(stream)>>>(wref, wtar, vol, out, W, H, D, r, dir);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_find_cross(const uint8_t *rgb, uint32_t *arms, int W, int H, int L, int tao, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size %dx%d", W, H);
    RTS_CHECK(L >= 0 && L <= 255, "arm limit %d must be in [0, 255]", L);
    dim3 grid(div_up(W, 64), div_up(H, 4)), block(64, 4);
    if (L <= FC_L)
        k_find_cross_t<false><<<dim3(div_up(W, 64), div_up(H, FC_TY)), block, 0, S(stream)>>>(rgb, arms, W, H, L, tao);
    else
        k_find_cross<<<grid, block, 0, S(stream)>>>(rgb, arms, W, H, L, tao);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_arms_unpack(const uint32_t *arms, int *horz, int *vert, int W, int H, void *stream) {
    const int P = W * H;
    k_arms_unpack<<<div_up(P, 256), 256, 0, S(stream)>>>(arms, horz, vert, P);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_arms_pack(const int *horz, const int *vert, uint32_t *arms, int W, int H, void *stream) {
    const int P = W * H;
    k_arms_pack<<<div_up(P, 256), 256, 0, S(stream)>>>(horz, vert, arms, P);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_hprefix(float *vol, int W, int H, int D, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    const long rows = (long)D * H;
    const int vec = (W % 4) == 0 && ((uintptr_t)vol % 16) == 0;
    k_hprefix<<<div_up(rows, PF_ROWS), 64, 0, S(stream)>>>(vol, rows, W, vec);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_cross_v(const float *P, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W, int H,
                       int D, int mode, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(P != out, "cross aggregation is out of place");
    const int base_mode = mode & ~RTS_CROSS_ARMS_LE17;
    RTS_CHECK(base_mode == RTS_CROSS_EXACT || base_mode == RTS_CROSS_FAST, "bad cross mode %d", mode);
    const int nx = div_up(W, CV_TX), nseg = div_up(H, SW_SEG);
    RTS_CHECK((long)nx * nseg * D < (1L << 31), "volume too large");
    // Both mode values run the same sweep, which adds in the reference's order.
    if (mode & RTS_CROSS_ARMS_LE17) {
        const int nsg = xs_segments(W, H, D), seg = div_up(H, nsg);
        RTS_CHECK((long)nx * nsg * D < (1L << 31), "volume too large");
        k_cross_sweep<XS_CH, XS_NW><<<nx * nsg * D, 64 * XS_NW, 0, S(stream)>>>(P, out, arms_ref, arms_tar, W, H, nx,
                                                                               nsg, seg, D, XS_GROUP);
    } else {
        k_cross_strip<<<nx * nseg * D, 256, 0, S(stream)>>>(P, out, arms_ref, arms_tar, W, H, nx, nseg, D);
    }
    RTS_LAUNCHED();
    return RTS_OK;
}

size_t rts_cross_v_wta_ws_bytes(int W, int H) {
    return W < 1 || H < 1 ? 0 : (size_t)W * (size_t)H * sizeof(unsigned long long);
}

rts_status rts_cross_v_wta(const float *P, const uint32_t *arms_ref, const uint32_t *arms_tar, uint8_t *dl,
                           uint8_t *dr, uint8_t *mask, int W, int H, int D, int tol, void *ws, size_t ws_bytes,
                           void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(dl != nullptr, "rts_cross_v_wta needs dl");
    RTS_CHECK(mask == nullptr || dr != nullptr, "the left-right mask needs dr");
    const size_t need = rts_cross_v_wta_ws_bytes(W, H);
    RTS_CHECK(dr == nullptr || (ws != nullptr && ws_bytes >= need && ((uintptr_t)ws & 7) == 0),
              "right-view workspace of %zu bytes (8-byte aligned) needed, %zu given", need, ws_bytes);
    const int nx = div_up(W, CV_TX), nseg = div_up(H, FW_SEG);
    hipStream_t s = S(stream);
    const size_t n = (size_t)W * H;
    auto *r2l = static_cast<unsigned long long *>(ws);
    if (dr) {
        k_fill_u64<<<(unsigned)div_up((long)n, 256L), 256, 0, s>>>(r2l, ~0ull, n);
        RTS_LAUNCHED();
        k_cross_sweep_wta<true><<<nx * nseg, 64 * FW_NW, 0, s>>>(P, arms_ref, arms_tar, dl, r2l, W, H, D, nx, nseg);
        RTS_LAUNCHED();
        k_r2l_finish<<<(unsigned)div_up((long)n, 256L), 256, 0, s>>>(r2l, dr, n);
        RTS_LAUNCHED();
        if (mask)
            if (int e = rts_cross_check(dl, dr, mask, W, H, tol, stream)) return e;
    } else {
        k_cross_sweep_wta<false><<<nx * nseg, 64 * FW_NW, 0, s>>>(P, arms_ref, arms_tar, dl, nullptr, W, H, D, nx,
                                                                  nseg);
        RTS_LAUNCHED();
    }
    return RTS_OK;
}

}  // extern "C"
