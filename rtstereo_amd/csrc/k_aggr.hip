// Cost aggregation kernels: box mean, adaptive weights, cross arms, the
// serial row prefix and the vertical cross sweep.  All volumes are
// disparity-major [D][H][W].
#include "rts_internal.h"

#pragma clang fp contract(off)

namespace rts {

// ---------------------------------------------------------------- box

// seq/ConstBlockAggregation.cpp:11-34, out of place.  One block = one d
// plane tile of 64 x BOX_TY outputs; the (64+2r) x (BOX_TY+2r) input tile is
// staged in LDS and every output sums its window in row-major order.
constexpr int BOX_TY = 16;
constexpr int BOX_RMAX = 16;

__global__ void __launch_bounds__(256) k_box_f32(const float *__restrict__ in, float *__restrict__ out, int W, int H,
                                                 int r) {
    __shared__ float tile[BOX_TY + 2 * BOX_RMAX][64 + 2 * BOX_RMAX + 1];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * BOX_TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const int tw = 64 + 2 * r, th = BOX_TY + 2 * r;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    for (int i = tid; i < tw * th; i += 256) {
        const int ty = i / tw, tx = i - ty * tw;
        const int gy = wrapn(y0 - r + ty, H), gx = wrapn(x0 - r + tx, W);
        tile[ty][tx] = in[base + (size_t)gy * W + gx];
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
        out[base + (size_t)y * W + x] = __fdiv_rn(acc, area);
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
            const int sub = lane >> 4, q = (lane & 15) * 4;
            for (int i = sub; i < nr; i += 4) {
                const float4 v = *reinterpret_cast<const float4 *>(vol + (r0 + i) * (long)W + c0 + q);
                tile[i][q] = v.x;
                tile[i][q + 1] = v.y;
                tile[i][q + 2] = v.z;
                tile[i][q + 3] = v.w;
            }
        } else {
            for (int i = 0; i < nr; i++)
                if (lane < nc) tile[i][lane] = vol[(r0 + i) * (long)W + c0 + lane];
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
// of CV_TX columns x CV_TY rows.  Rows y0-LMAX .. y0+CV_TY+LMAX-1 of the
// horizontal segment sums S(row, x) = P[row][x+r] - P[row][max(x-l-1, 0)] and
// their lengths l+r+1 are built once in LDS (l, r = min of the ref arm at
// (row, x) and the tar arm at (row, x-d)); each output then adds its rows.
// EXACT adds S in the reference's row order in float; FAST prefix-sums S
// down each column in FP64 and takes one difference.
constexpr int CV_TX = 64;
constexpr int CV_TY = 32;
constexpr int CV_LMAX = 17;   // tile halo: the reference's arm limit L (longer arms take a slow path)

template <bool EXACT>
__global__ void __launch_bounds__(256) k_cross_v(const float *__restrict__ P, float *__restrict__ out,
                                                 const uint32_t *__restrict__ aref, const uint32_t *__restrict__ atar,
                                                 int W, int H, int Lmax) {
    constexpr int RMAX = CV_TY + 2 * CV_LMAX;
    __shared__ float seg[EXACT ? RMAX : 1][CV_TX];
    __shared__ double cum[EXACT ? 1 : RMAX + 1][CV_TX];
    __shared__ int len[RMAX + 1][CV_TX];
    const int x0 = blockIdx.x * CV_TX, y0 = blockIdx.y * CV_TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const int rlo = max(0, y0 - Lmax), rhi = min(H, y0 + CV_TY + Lmax);   // rows [rlo, rhi)
    const int nrows = rhi - rlo;
    const int lx = threadIdx.x & 63, lg = threadIdx.x >> 6;   // column, row group (4)
    const int x = x0 + lx;
    const bool colok = x < W;
    const int tx = colok ? wrap1(x - d, W) : 0;
    // Segment sums per (row, column).
    for (int i = lg; i < nrows; i += 4) {
        float s = 0.0f;
        int n = 0;
        if (colok) {
            const int row = rlo + i;
            const uint32_t ar = aref[(size_t)row * W + x], at = atar[(size_t)row * W + tx];
            const int l = min(arm_l(ar), arm_l(at)), r = min(arm_r(ar), arm_r(at));
            // x-l-1 < 0 clamps to column 0 (the reference then drops column
            // 0's cost, seq/CrossBasedAggregation.cpp:151-156); x+r never
            // leaves the image for arms from findCross.
            const int xl = max(x - l - 1, 0), xr = min(x + r, W - 1);
            const float *Pr = P + base + (size_t)row * W;
            s = Pr[xr] - Pr[xl];
            n = l + r + 1;
        }
        if (EXACT) {
            seg[i][lx] = s;
            len[i][lx] = n;
        } else {
            cum[i + 1][lx] = (double)s;
            len[i + 1][lx] = n;
        }
    }
    __syncthreads();
    if (!EXACT) {
        // Column prefix sums (FP64 segments, int lengths); one wave per 16 columns.
        if (threadIdx.x < CV_TX) {
            double a = 0.0;
            int c = 0;
            cum[0][lx] = 0.0;
            len[0][lx] = 0;
            for (int i = 1; i <= nrows; i++) {
                a += cum[i][lx];
                c += len[i][lx];
                cum[i][lx] = a;
                len[i][lx] = c;
            }
        }
        __syncthreads();
    }
    if (!colok) return;
    for (int ly = lg; ly < CV_TY; ly += 4) {
        const int y = y0 + ly;
        if (y >= H) break;
        const uint32_t ar = aref[(size_t)y * W + x], at = atar[(size_t)y * W + tx];
        const int t = min(arm_t(ar), arm_t(at)), b = min(arm_b(ar), arm_b(at));
        const int ia = y - t - rlo, ib = y + b - rlo;   // inclusive row range in the tile
        float v;
        if (ia < 0 || ib >= nrows) {
            // Arms longer than the tile halo (never produced by rts_find_cross
            // with L <= CV_LMAX): sum the rows straight from global memory,
            // skipping rows outside the image as the reference does.
            float sum = 0.0f;
            int cnt = 0;
            for (int row = max(y - t, 0); row <= min(y + b, H - 1); row++) {
                const uint32_t rr = aref[(size_t)row * W + x], rt = atar[(size_t)row * W + tx];
                const int l = min(arm_l(rr), arm_l(rt)), r = min(arm_r(rr), arm_r(rt));
                const int xl = max(x - l - 1, 0), xr = min(x + r, W - 1);
                const float *Pr = P + base + (size_t)row * W;
                sum += Pr[xr] - Pr[xl];
                cnt += l + r + 1;
            }
            out[base + (size_t)y * W + x] = __fdiv_rn(sum, (float)cnt);
            continue;
        }
        if (EXACT) {
            float sum = 0.0f;
            int cnt = 0;
            for (int i = ia; i <= ib; i++) {
                sum += seg[i][lx];
                cnt += len[i][lx];
            }
            v = __fdiv_rn(sum, (float)cnt);
        } else {
            const double sum = cum[ib + 1][lx] - cum[ia][lx];
            const int cnt = len[ib + 1][lx] - len[ia][lx];
            v = (float)(sum / (double)cnt);
        }
        out[base + (size_t)y * W + x] = v;
    }
}

// One whole cross-based aggregation from a raw cost volume.  The block's
// cost tile (rows y0-L .. y0+CV_TY+L-1, columns x0-L-1 .. x0+CV_TX+L-1,
// clipped to the image) is prefix-summed along each row in LDS, starting at
// the tile's first column; segment sums are differences of that local
// prefix (the same sums as the reference's global prefix, including the
// column-0 drop when the tile starts at column 0), then prefix-summed down
// each column in FP64 as the FAST sweep does.  Integer-valued costs stay
// exact throughout.
constexpr int CF_COLS = CV_TX + 2 * CV_LMAX + 1;
constexpr int CF_ROWS = CV_TY + 2 * CV_LMAX;

__global__ void __launch_bounds__(256) k_cross_fused(const float *__restrict__ C, float *__restrict__ out,
                                                     const uint32_t *__restrict__ aref,
                                                     const uint32_t *__restrict__ atar, int W, int H) {
    __shared__ float tile[CF_ROWS][CF_COLS + 1];
    __shared__ double cum[CF_ROWS + 1][CV_TX];
    __shared__ short len[CF_ROWS + 1][CV_TX];
    const int x0 = blockIdx.x * CV_TX, y0 = blockIdx.y * CV_TY, d = blockIdx.z;
    const size_t base = (size_t)d * plane(W, H);
    const int rlo = max(0, y0 - CV_LMAX), rhi = min(H, y0 + CV_TY + CV_LMAX);
    const int clo = max(0, x0 - CV_LMAX - 1), chi = min(W, x0 + CV_TX + CV_LMAX);
    const int nrows = rhi - rlo, ncols = chi - clo;
    for (int i = threadIdx.x; i < nrows * ncols; i += 256) {
        const int rr = i / ncols, cc = i - rr * ncols;
        tile[rr][cc] = C[base + (size_t)(rlo + rr) * W + clo + cc];
    }
    __syncthreads();
    for (int rr = threadIdx.x; rr < nrows; rr += 256) {
        float a = tile[rr][0];
        for (int cc = 1; cc < ncols; cc++) {
            a = a + tile[rr][cc];
            tile[rr][cc] = a;
        }
    }
    __syncthreads();
    const int lx = threadIdx.x & 63, lg = threadIdx.x >> 6;
    const int x = x0 + lx;
    const bool colok = x < W;
    const int tx = colok ? wrap1(x - d, W) : 0;
    for (int i = lg; i < nrows; i += 4) {
        double s = 0.0;
        int n = 0;
        if (colok) {
            const int row = rlo + i;
            const uint32_t ar = aref[(size_t)row * W + x], at = atar[(size_t)row * W + tx];
            const int l = min(min(arm_l(ar), arm_l(at)), CV_LMAX), r = min(min(arm_r(ar), arm_r(at)), CV_LMAX);
            const int xl = max(x - l - 1, 0) - clo, xr = min(x + r, W - 1) - clo;
            s = (double)(tile[i][xr] - tile[i][xl]);
            n = l + r + 1;
        }
        cum[i + 1][lx] = s;
        len[i + 1][lx] = (short)n;
    }
    __syncthreads();
    if (threadIdx.x < CV_TX) {
        double a = 0.0;
        int c = 0;
        cum[0][lx] = 0.0;
        len[0][lx] = 0;
        for (int i = 1; i <= nrows; i++) {
            a += cum[i][lx];
            c += len[i][lx];
            cum[i][lx] = a;
            len[i][lx] = (short)c;
        }
    }
    __syncthreads();
    if (!colok) return;
    for (int ly = lg; ly < CV_TY; ly += 4) {
        const int y = y0 + ly;
        if (y >= H) break;
        const uint32_t ar = aref[(size_t)y * W + x], at = atar[(size_t)y * W + tx];
        const int t = min(arm_t(ar), arm_t(at)), b = min(arm_b(ar), arm_b(at));
        const int ia = max(y - t, rlo) - rlo, ib = min(y + b, rhi - 1) - rlo;
        const double sum = cum[ib + 1][lx] - cum[ia][lx];
        const int cnt = len[ib + 1][lx] - len[ia][lx];
        out[base + (size_t)y * W + x] = (float)(sum / (double)cnt);
    }
}

}  // namespace rts

using namespace rts;

extern "C" {

rts_status rts_cross_fused(const float *cost, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W,
                           int H, int D, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(cost != out, "cross aggregation is out of place");
    RTS_CHECK(div_up(H, CV_TY) <= 65535, "height %d too large", H);
    dim3 grid(div_up(W, CV_TX), div_up(H, CV_TY), D);
    k_cross_fused<<<grid, 256, 0, S(stream)>>>(cost, out, arms_ref, arms_tar, W, H);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_box(const float *in, float *out, int W, int H, int D, int r, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r < W && r < H, "box radius %d does not fit the image", r);
    RTS_CHECK(in != out, "rts_box is out of place");
    if (r <= BOX_RMAX) {
        dim3 grid(div_up(W, 64), div_up(H, BOX_TY), D), block(64, 4);
        k_box_f32<<<grid, block, 0, S(stream)>>>(in, out, W, H, r);
    } else {
        dim3 grid(div_up(W, 64), div_up(H, 4), D), block(64, 4);
        k_box_f32_global<<<grid, block, 0, S(stream)>>>(in, out, W, H, r);
    }
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

// Largest arm in either map bounds the tile halo; arms above CV_LMAX are
// rejected (the reference's arms are at most 17).
rts_status rts_cross_v(const float *P, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W, int H,
                       int D, int mode, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(P != out, "cross aggregation is out of place");
    RTS_CHECK(mode == RTS_CROSS_EXACT || mode == RTS_CROSS_FAST, "bad cross mode %d", mode);
    RTS_CHECK(div_up(H, CV_TY) <= 65535, "height %d too large", H);
    dim3 grid(div_up(W, CV_TX), div_up(H, CV_TY), D);
    if (mode == RTS_CROSS_EXACT)
        k_cross_v<true><<<grid, 256, 0, S(stream)>>>(P, out, arms_ref, arms_tar, W, H, CV_LMAX);
    else
        k_cross_v<false><<<grid, 256, 0, S(stream)>>>(P, out, arms_ref, arms_tar, W, H, CV_LMAX);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
