// Adaptive-weight aggregation (seq/AdaptiveWeightAggregation.cpp:11-111)
// without the reference's per-pixel weight tensors over the whole image.
//
// For pixel p = (x, y), disparity d and target centre q = (x-d wrapped, y):
//   w_ref(p,k) = prox(k) * exp(-|Lab_L(p) - Lab_L(p+k)| / gc)   (0 outside)
//   w_tar(q,k) = prox(k) * exp(-|Lab_R(q) - Lab_R(q+k)| / gc)   (0 outside)
//   out(p,d)   = sum_k w_ref*w_tar*C(p+k wrapped, d) / sum_k w_ref*w_tar
//
// The image is processed in bands of rows.  For each band a first kernel
// writes both views' weights of the band's centre pixels to a workspace,
// laid out [row][dy][tap][x] so that a tap row of consecutive pixels is
// contiguous.  The aggregation kernel then works on tiles of 64 pixels x 2
// rows x 64 disparities: lane = disparity, wave = 8 consecutive pixels.
// Per window row it stages in LDS the 64 cost rows of its disparities and
// the target weights of the 127 centres x-d the tile touches; the reference
// weights of a wave's 8 pixels are uniform across its lanes and come in
// through scalar loads.  Each lane then accumulates 2 rows x 8 pixels of
// its disparity, reading the target weight from LDS and the cost from a
// register window that slides along the tap row.
//
//   FP32: product, sum, fused multiply-add per tap (3 VALU);
//   FP16: weights stored as half pairs of neighbouring taps, one packed
//         product per pair and two v_dot2_f32_f16 (numerator, denominator)
//         accumulating in float (1.5 VALU per tap).
#include "rts_internal.h"

#include <hip/hip_fp16.h>

namespace rts {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));

constexpr int AW_RMAX = 16;               // largest window radius
constexpr int AW_TX = 64;                 // pixels per tile
constexpr int AW_NX = 8;                  // pixels per lane
constexpr int AW_NY = 2;                  // rows per tile
constexpr size_t AW_WS_CAP = size_t(1) << 30;   // workspace bytes for both views

// Weight entries per (pixel, window row): tap pairs in FP16, taps in FP32.
__host__ __device__ inline int aw_ntap(int R, bool half) { return half ? R + 1 : 2 * R + 1; }

__host__ inline int aw_rmax_for(int r) { return r <= 3 ? 3 : r <= 5 ? 5 : r <= 8 ? 8 : 16; }

// Weights of band rows [y0, y0+rows) of one view: wb[ry][dy+r][t][x], x < Wp.
// Block: 64 pixels of one (row, window row); the four waves split the taps.
// prox(dy, dx) = exp(-sqrt(dy^2+dx^2)/gp) comes from a per-block table; the
// colour term is exp2(|dLab| * kc) with kc = -log2(e)/gc (hardware exp2 and
// square root, each within an ulp or two: relative weight error ~1e-6).
template <bool HALF, int R>
__global__ void __launch_bounds__(256) k_aw_weights_band(const uint8_t *__restrict__ lab, uint32_t *__restrict__ wb,
                                                         int W, int H, int Wp, int y0, int r, float kc, float gp) {
    constexpr int NT = HALF ? R + 1 : 2 * R + 1;
    __shared__ float prox[2 * R + 2];
    const int e = 2 * r + 1;
    const int row = blockIdx.y, ry = row / e, dy = row - ry * e - r;
    const int y = y0 + ry, yy = y + dy;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    if (threadIdx.x < 2 * R + 2) {
        const int dx = (int)threadIdx.x - R;
        prox[threadIdx.x] =
            (dx > r || dx < -r || dy > r || dy < -r) ? 0.0f : expf(-sqrtf((float)(dy * dy + dx * dx)) / gp);
    }
    __syncthreads();
    const bool ok = x < W && yy >= 0 && yy < H;
    float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
    const uint8_t *nrow = lab + 3 * (size_t)(ok ? yy : 0) * W;
    if (ok) {
        const uint8_t *c = lab + 3 * ((size_t)y * W + x);
        c0 = c[0];
        c1 = c[1];
        c2 = c[2];
    }
    auto weight = [&](int dx) {
        const int xx = x + dx;
        if (!ok || xx < 0 || xx >= W) return 0.0f;
        const uint8_t *t = nrow + 3 * xx;
        const float e0 = c0 - (float)t[0], e1 = c1 - (float)t[1], e2 = c2 - (float)t[2];
        const float dist = __builtin_amdgcn_sqrtf(e0 * e0 + e1 * e1 + e2 * e2);
        return prox[dx + R] * __builtin_amdgcn_exp2f(dist * kc);
    };
    uint32_t *dst = wb + (size_t)row * NT * Wp + x;
    for (int jt = threadIdx.x >> 6; jt < NT; jt += 4) {
        if (HALF) {
            const int dx = 2 * jt - R;
            h2 v;
            v.x = (_Float16)weight(dx);
            v.y = (_Float16)weight(dx + 1);
            dst[(size_t)jt * Wp] = __builtin_bit_cast(uint32_t, v);
        } else {
            dst[(size_t)jt * Wp] = __float_as_uint(weight(jt - R));
        }
    }
}

__device__ __forceinline__ int wrap_any(int v, int n) {
    if (v < 0) v += n;
    else if (v >= n) v -= n;
    return (v < 0 || v >= n) ? wrapn(v, n) : v;
}

// LDS layout and staging shape of one aggregation tile.
template <bool HALF, int R>
struct AwTile {
    static constexpr int NT = HALF ? R + 1 : 2 * R + 1;   // weight words per (pixel, window row)
    // Weight rows padded to 16 bytes and read with ds_read_b128.  A pitch of
    // 20 or 36 words puts 16 consecutive lanes (consecutive centres) on 16
    // distinct 4-bank groups.
    static constexpr int PJ = (NT + 3) & ~3;
    static constexpr int PJW = PJ;
    static constexpr int CQ = (AW_TX + 2 * R + 2 + 3) / 4; // cost quads per disparity row
    static constexpr int PC = HALF ? (2 * CQ) | 1 : (4 * CQ) | 1;   // cost row pitch in words (odd)
    static constexpr int C_WORDS = (64 * PC + 3) & ~3;
    static constexpr int T_WORDS = AW_NY * 128 * PJ;
    static constexpr int L_WORDS = AW_NY * 64 * PJW;
    static constexpr int WORDS = (C_WORDS + T_WORDS + L_WORDS) > 64 * 65 ? C_WORDS + T_WORDS + L_WORDS : 64 * 65;
    static constexpr int SQ_C = 4;                                   // 64 rows x 32 quad slots / 512
    static constexpr int SQ_T = (AW_NY * NT * 32 + 511) / 512;       // target weights: 128 centres per tap word
    static constexpr int SQ_L = (AW_NY * NT * 16 + 511) / 512;       // reference weights: 64 pixels per tap word
};

// One window row's worth of tile inputs, held in registers between the
// global loads and the LDS stores.
template <bool HALF, int R>
struct AwStage {
    using T = AwTile<HALF, R>;
    uint4 c[T::SQ_C], t[T::SQ_T], l[T::SQ_L];
};

template <bool HALF, int R>
__global__ void __launch_bounds__(512) k_aw_band(const uint32_t *__restrict__ wlb, const uint32_t *__restrict__ wtb,
                                                 const float *__restrict__ vol, float *__restrict__ out, int W, int H,
                                                 int D, int Wp, int y0, int rows, int r, int nxb, int nyb, int ndb) {
    using T = AwTile<HALF, R>;
    constexpr int NT = T::NT;
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    uint32_t *Ct = lds, *Tt = lds + T::C_WORDS, *Lt = lds + T::C_WORDS + T::T_WORDS;

    const int nblk = nxb * nyb * ndb;
    const int lin = xcd_order(blockIdx.x, nblk);
    const int xb = lin % nxb, rest = lin / nxb;
    const int db = rest % ndb, yb = rest / ndb;
    const int x0 = xb * AW_TX, d0 = db * 64, ry0 = yb * AW_NY, yt = y0 + ry0;
    const int tid = threadIdx.x, dl = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int bx = wave * AW_NX;
    const int nrows = min(AW_NY, rows - ry0);
    const int e = 2 * r + 1;
    const size_t pl = plane(W, H);
    const int q_lo = x0 - d0 - 64;   // first target centre staged
    const bool fast_c = (W & 3) == 0 && ((x0 - R) & 3) == 0 && x0 - R >= 0 && x0 - R + 4 * T::CQ <= W;
    const bool fast_t = q_lo >= 0 && q_lo + 128 <= W;

    auto row_valid = [&](int yy, int ny) {
        const int dy = yy - (yt + ny);
        return ny < nrows && dy >= -r && dy <= r;
    };
    auto wrow = [&](const uint32_t *base, int yy, int ny, int j) {
        return base + ((size_t)(ry0 + ny) * e + (yy - (yt + ny) + r)) * NT * Wp + (size_t)j * Wp;
    };

    auto load = [&](AwStage<HALF, R> &st, int yy) {
        const float *crow = vol + (size_t)wrap_any(yy, H) * W;
#pragma unroll
        for (int s = 0; s < T::SQ_C; s++) {
            const int dd = (wave << 3) + (s << 1) + (dl >> 5), qc = dl & 31, d = d0 + dd;
            uint4 v = {0u, 0u, 0u, 0u};
            if (qc < T::CQ && d < D) {
                const float *src = crow + (size_t)d * pl;
                const int c0 = x0 - R + 4 * qc;
                if (fast_c) {
                    v = *reinterpret_cast<const uint4 *>(src + c0);
                } else {
                    v.x = __float_as_uint(src[wrap_any(c0, W)]);
                    v.y = __float_as_uint(src[wrap_any(c0 + 1, W)]);
                    v.z = __float_as_uint(src[wrap_any(c0 + 2, W)]);
                    v.w = __float_as_uint(src[wrap_any(c0 + 3, W)]);
                }
            }
            st.c[s] = v;
        }
#pragma unroll
        for (int s = 0; s < T::SQ_T; s++) {
            const int k = s * 512 + tid;
            const int qq = k & 31, jj = k >> 5, ny = jj >= NT ? 1 : 0, j = jj - ny * NT;
            uint4 v = {0u, 0u, 0u, 0u};
            if (k < AW_NY * NT * 32 && row_valid(yy, ny)) {
                const uint32_t *src = wrow(wtb, yy, ny, j);
                const int q = q_lo + 4 * qq;
                if (fast_t) {
                    v = *reinterpret_cast<const uint4 *>(src + q);
                } else {
                    v.x = src[wrap_any(q, W)];
                    v.y = src[wrap_any(q + 1, W)];
                    v.z = src[wrap_any(q + 2, W)];
                    v.w = src[wrap_any(q + 3, W)];
                }
            }
            st.t[s] = v;
        }
#pragma unroll
        for (int s = 0; s < T::SQ_L; s++) {
            const int k = s * 512 + tid;
            const int pq = k & 15, jj = k >> 4, ny = jj >= NT ? 1 : 0, j = jj - ny * NT;
            uint4 v = {0u, 0u, 0u, 0u};
            if (k < AW_NY * NT * 16 && row_valid(yy, ny))
                v = *reinterpret_cast<const uint4 *>(wrow(wlb, yy, ny, j) + x0 + 4 * pq);
            st.l[s] = v;
        }
    };

    auto store = [&](const AwStage<HALF, R> &st) {
#pragma unroll
        for (int s = 0; s < T::SQ_C; s++) {
            const int dd = (wave << 3) + (s << 1) + (dl >> 5), qc = dl & 31;
            if (qc >= T::CQ) continue;
            const uint4 v = st.c[s];
            if (HALF) {
                h2 a, b;
                a.x = (_Float16)__uint_as_float(v.x);
                a.y = (_Float16)__uint_as_float(v.y);
                b.x = (_Float16)__uint_as_float(v.z);
                b.y = (_Float16)__uint_as_float(v.w);
                Ct[dd * T::PC + 2 * qc] = __builtin_bit_cast(uint32_t, a);
                Ct[dd * T::PC + 2 * qc + 1] = __builtin_bit_cast(uint32_t, b);
            } else {
                uint32_t *dst = Ct + dd * T::PC + 4 * qc;
                dst[0] = v.x;
                dst[1] = v.y;
                dst[2] = v.z;
                dst[3] = v.w;
            }
        }
#pragma unroll
        for (int s = 0; s < T::SQ_T; s++) {
            const int k = s * 512 + tid;
            if (k >= AW_NY * NT * 32) continue;
            const int qq = k & 31, jj = k >> 5, ny = jj >= NT ? 1 : 0, j = jj - ny * NT;
            uint32_t *dst = Tt + (ny * 128 + 4 * qq) * T::PJ + j;
            dst[0] = st.t[s].x;
            dst[T::PJ] = st.t[s].y;
            dst[2 * T::PJ] = st.t[s].z;
            dst[3 * T::PJ] = st.t[s].w;
        }
#pragma unroll
        for (int s = 0; s < T::SQ_L; s++) {
            const int k = s * 512 + tid;
            if (k >= AW_NY * NT * 16) continue;
            const int pq = k & 15, jj = k >> 4, ny = jj >= NT ? 1 : 0, j = jj - ny * NT;
            uint32_t *dst = Lt + (ny * 64 + 4 * pq) * T::PJW + j;
            dst[0] = st.l[s].x;
            dst[T::PJW] = st.l[s].y;
            dst[2 * T::PJW] = st.l[s].z;
            dst[3 * T::PJW] = st.l[s].w;
        }
    };

    float num[AW_NY][AW_NX], den[AW_NY][AW_NX];
#pragma unroll
    for (int a = 0; a < AW_NY; a++)
#pragma unroll
        for (int m = 0; m < AW_NX; m++) num[a][m] = den[a][m] = 0.0f;

    auto compute = [&](int yy) {
        if (HALF) {
            h2 ce[AW_NX / 2 + R + 1], co[AW_NX / 2 + R];
            const uint32_t *crow = Ct + dl * T::PC + bx / 2;
#pragma unroll
            for (int i = 0; i < AW_NX / 2 + R + 1; i++) ce[i] = __builtin_bit_cast(h2, crow[i]);
#pragma unroll
            for (int i = 0; i < AW_NX / 2 + R; i++) co[i] = __builtin_shufflevector(ce[i], ce[i + 1], 1, 2);
            const h2 one = {(_Float16)1.0f, (_Float16)1.0f};
#pragma unroll
            for (int ny = 0; ny < AW_NY; ny++) {
                if (!row_valid(yy, ny)) continue;
                const uint4 *wt = reinterpret_cast<const uint4 *>(Tt + (ny * 128 + bx - dl + 64) * T::PJ);
                const uint4 *wl = reinterpret_cast<const uint4 *>(Lt + (ny * 64 + bx) * T::PJW);
#pragma unroll
                for (int m = 0; m < AW_NX; m++) {
                    uint4 tv[T::PJ / 4], lv[T::PJ / 4];
#pragma unroll
                    for (int k = 0; k < T::PJ / 4; k++) {
                        tv[k] = wt[m * (T::PJ / 4) + k];
                        lv[k] = wl[m * (T::PJW / 4) + k];
                    }
                    const uint32_t *tw = reinterpret_cast<const uint32_t *>(tv);
                    const uint32_t *lw = reinterpret_cast<const uint32_t *>(lv);
                    // All products first: a dot2 reading a product written by
                    // the instruction just before it costs a wait state.
                    h2 p[NT];
#pragma unroll
                    for (int j = 0; j < NT; j++) p[j] = __builtin_bit_cast(h2, lw[j]) * __builtin_bit_cast(h2, tw[j]);
                    float sn = num[ny][m], sd = den[ny][m];
#pragma unroll
                    for (int j = 0; j < NT; j++) {
                        const h2 c = (m & 1) ? co[(m >> 1) + j] : ce[(m >> 1) + j];
                        sn = __builtin_amdgcn_fdot2(p[j], c, sn, false);
                        sd = __builtin_amdgcn_fdot2(p[j], one, sd, false);
                    }
                    num[ny][m] = sn;
                    den[ny][m] = sd;
                }
            }
        } else {
            float cf[AW_NX + 2 * R];
            const float *crow = reinterpret_cast<const float *>(Ct) + dl * T::PC + bx;
#pragma unroll
            for (int i = 0; i < AW_NX + 2 * R; i++) cf[i] = crow[i];
#pragma unroll
            for (int ny = 0; ny < AW_NY; ny++) {
                if (!row_valid(yy, ny)) continue;
                const float4 *wt = reinterpret_cast<const float4 *>(Tt + (ny * 128 + bx - dl + 64) * T::PJ);
                const float4 *wl = reinterpret_cast<const float4 *>(Lt + (ny * 64 + bx) * T::PJW);
#pragma unroll
                for (int m = 0; m < AW_NX; m++) {
                    float4 tv[T::PJ / 4], lv[T::PJ / 4];
#pragma unroll
                    for (int k = 0; k < T::PJ / 4; k++) {
                        tv[k] = wt[m * (T::PJ / 4) + k];
                        lv[k] = wl[m * (T::PJW / 4) + k];
                    }
                    const float *tw = reinterpret_cast<const float *>(tv);
                    const float *lw = reinterpret_cast<const float *>(lv);
                    float sn = num[ny][m], sd = den[ny][m];
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        const float p = lw[t] * tw[t];
                        sd += p;
                        sn = __builtin_fmaf(p, cf[m + t], sn);
                    }
                    num[ny][m] = sn;
                    den[ny][m] = sd;
                }
            }
        }
    };

    const int ylast = yt + nrows - 1 + r;
    AwStage<HALF, R> st;
    if (HALF) load(st, yt - r);
    for (int yy = yt - r; yy <= ylast; yy++) {
        __syncthreads();
        if (!HALF) load(st, yy);
        store(st);
        __syncthreads();
        if (HALF && yy < ylast) load(st, yy + 1);   // in flight during the arithmetic
        compute(yy);
    }

    // Write through LDS so that each disparity row of 64 pixels is stored
    // contiguously.
    float *tile = reinterpret_cast<float *>(lds);
#pragma unroll
    for (int ny = 0; ny < AW_NY; ny++) {
        if (ny >= nrows) break;
        __syncthreads();
#pragma unroll
        for (int m = 0; m < AW_NX; m++) tile[dl * 65 + bx + m] = num[ny][m] / den[ny][m];
        __syncthreads();
        const int y = yt + ny;
        for (int i = tid; i < 64 * AW_TX; i += 512) {
            const int dd = i >> 6, px = i & 63;
            const int d = d0 + dd, x = x0 + px;
            if (d < D && x < W) out[(size_t)d * pl + (size_t)y * W + x] = tile[dd * 65 + px];
        }
    }
}

template <bool HALF, int R>
int aw_band_launch(const void *wl, const void *wt, const float *vol, float *out, int W, int H, int D, int Wp, int y0,
                   int rows, int r, hipStream_t s) {
    using T = AwTile<HALF, R>;
    constexpr size_t lds_bytes = sizeof(uint32_t) * T::WORDS;
    static bool attr_set = false;   // the attribute is per function; setting it twice is harmless
    if (lds_bytes > 65536 && !attr_set) {
        RTS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_aw_band<HALF, R>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
        attr_set = true;
    }
    const int nxb = div_up(W, AW_TX), nyb = div_up(rows, AW_NY), ndb = div_up(D, 64);
    k_aw_band<HALF, R><<<nxb * nyb * ndb, 512, lds_bytes, s>>>(static_cast<const uint32_t *>(wl),
                                                               static_cast<const uint32_t *>(wt), vol, out, W, H, D,
                                                               Wp, y0, rows, r, nxb, nyb, ndb);
    RTS_LAUNCHED();
    return RTS_OK;
}

struct AwPlan {
    int R, NT, Wp, band, ntap_bytes;
    size_t view_bytes;   // workspace of one view for one band
};

AwPlan aw_plan(int W, int H, int r, bool half) {
    AwPlan p;
    p.R = aw_rmax_for(r);
    p.NT = aw_ntap(p.R, half);
    p.Wp = div_up(W, AW_TX) * AW_TX;
    const size_t row = (size_t)(2 * r + 1) * p.NT * p.Wp * 4;   // one view, one image row
    int band = (int)std::min<size_t>((size_t)H, std::max<size_t>(AW_NY, AW_WS_CAP / (2 * row)));
    band = std::min(band, 65535 / (2 * r + 1));   // grid.y of the weight kernel
    band -= band % AW_NY;
    if (band < AW_NY) band = AW_NY;
    if (band > H) band = H;
    p.band = band;
    p.view_bytes = row * band;
    return p;
}

template <bool HALF, int R>
int aw_run(const uint8_t *labL, const uint8_t *labR, const float *vol, float *out, int W, int H, int D, int r, float gc,
           float gp, const AwPlan &pl, uint8_t *ws, hipStream_t s) {
    void *wl = ws, *wt = ws + pl.view_bytes;
    for (int y0 = 0; y0 < H; y0 += pl.band) {
        const int rows = std::min(pl.band, H - y0);
        const dim3 grid(pl.Wp / 64, rows * (2 * r + 1));
        const float kc = (float)(-1.4426950408889634 / (double)gc);
        k_aw_weights_band<HALF, R><<<grid, 256, 0, s>>>(labL, static_cast<uint32_t *>(wl), W, H, pl.Wp, y0, r, kc, gp);
        RTS_LAUNCHED();
        k_aw_weights_band<HALF, R><<<grid, 256, 0, s>>>(labR, static_cast<uint32_t *>(wt), W, H, pl.Wp, y0, r, kc, gp);
        RTS_LAUNCHED();
        if (int st = aw_band_launch<HALF, R>(wl, wt, vol, out, W, H, D, pl.Wp, y0, rows, r, s)) return st;
    }
    return RTS_OK;
}

template <bool HALF>
int aw_dispatch(const uint8_t *labL, const uint8_t *labR, const float *vol, float *out, int W, int H, int D, int r,
                float gc, float gp, const AwPlan &pl, uint8_t *ws, hipStream_t s) {
    switch (pl.R) {
        case 3: return aw_run<HALF, 3>(labL, labR, vol, out, W, H, D, r, gc, gp, pl, ws, s);
        case 5: return aw_run<HALF, 5>(labL, labR, vol, out, W, H, D, r, gc, gp, pl, ws, s);
        case 8: return aw_run<HALF, 8>(labL, labR, vol, out, W, H, D, r, gc, gp, pl, ws, s);
        default: return aw_run<HALF, 16>(labL, labR, vol, out, W, H, D, r, gc, gp, pl, ws, s);
    }
}

}  // namespace rts

using namespace rts;

static int aw_check(const float *vol, float *out, int W, int H, int D, int r, int precision) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(r >= 0 && r <= AW_RMAX && r < W && r < H, "AW radius %d must be in [0, %d] and fit the image", r,
              AW_RMAX);
    RTS_CHECK(vol != out, "AW aggregation is out of place");
    RTS_CHECK(precision == RTS_AW_FP32 || precision == RTS_AW_FP16, "bad AW precision %d", precision);
    return RTS_OK;
}

extern "C" size_t rts_aw_workspace_bytes(int W, int H, int r, int precision) {
    if (W < 1 || H < 1 || r < 0 || r > AW_RMAX) return 0;
    return 2 * aw_plan(W, H, r, precision == RTS_AW_FP16).view_bytes;
}

extern "C" rts_status rts_aw_ws(const uint8_t *lab_ref, const uint8_t *lab_tar, const float *vol, float *out, int W,
                                int H, int D, int r, float gamma_c, float gamma_p, int precision, void *workspace,
                                size_t workspace_bytes, void *stream) {
    if (int e = aw_check(vol, out, W, H, D, r, precision)) return e;
    const bool half = precision == RTS_AW_FP16;
    const AwPlan pl = aw_plan(W, H, r, half);
    RTS_CHECK(workspace && workspace_bytes >= 2 * pl.view_bytes, "AW workspace of %zu bytes is below the %zu needed",
              workspace_bytes, 2 * pl.view_bytes);
    hipStream_t s = S(stream);
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    return half ? aw_dispatch<true>(lab_ref, lab_tar, vol, out, W, H, D, r, gamma_c, gamma_p, pl, ws, s)
                : aw_dispatch<false>(lab_ref, lab_tar, vol, out, W, H, D, r, gamma_c, gamma_p, pl, ws, s);
}

extern "C" rts_status rts_aw(const uint8_t *lab_ref, const uint8_t *lab_tar, const float *vol, float *out, int W,
                             int H, int D, int r, float gamma_c, float gamma_p, int precision, void *stream) {
    if (int e = aw_check(vol, out, W, H, D, r, precision)) return e;
    const size_t need = rts_aw_workspace_bytes(W, H, r, precision);
    hipStream_t s = S(stream);
    void *ws = nullptr;
    RTS_HIP(hipMallocAsync(&ws, need, s));
    const int st = rts_aw_ws(lab_ref, lab_tar, vol, out, W, H, D, r, gamma_c, gamma_p, precision, ws, need, stream);
    (void)hipFreeAsync(ws, s);
    return st;
}
