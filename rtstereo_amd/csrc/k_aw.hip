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
// rows x 64 disparities: lane = disparity, wave = 8 pixels (FP32: 8
// consecutive pixels; FP16: the even or the odd pixels of a group of 16).
// Per window row it stages in LDS the 64 cost rows of its disparities, the
// target weights of the 127 centres x-d the tile touches and the reference
// weights of the tile's pixels; a wave's reference weights are the same
// address on all its lanes (broadcast LDS reads).  Each lane then
// accumulates 2 rows x 8 pixels of its disparity, reading the weights from
// LDS and the cost from a register window that slides along the tap row.
//
//   FP32: product, sum, fused multiply-add per tap (3 VALU);
//   FP16: weights stored as half pairs of neighbouring taps, one packed
//         product per pair and two v_dot2_f32_f16 (numerator, denominator)
//         accumulating in float (1.5 VALU per tap).
//
// Two FP16 kernels follow the tile kernel: k_aw_band_dma (a float cost
// volume; one tile row per block, staging by direct loads into LDS, the
// denominators on the matrix cores, one VALU per tap) and k_aw_band_h (an
// FP16 cost volume and up to three disparities per lane, so that each weight
// row read from LDS serves several outputs), which give the same bits.
#include "rts_internal.h"

#include <hip/hip_fp16.h>

#include <type_traits>

namespace rts {

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 half4 __attribute__((ext_vector_type(4)));
typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int AW_RMAX = 16;               // largest window radius
constexpr int AW_TX = 64;                 // weight rows are padded to a multiple of this
constexpr int AW_NX = 8;                  // pixels per lane
constexpr int AW_NY = 2;                  // rows per tile
constexpr size_t AW_WS_CAP = size_t(2) << 30;   // workspace bytes for both views

// Weight entries per (pixel, window row): tap pairs in FP16, taps in FP32.
__host__ __device__ inline int aw_ntap(int R, bool half) { return half ? R + 1 : 2 * R + 1; }

__host__ inline int aw_rmax_for(int r) { return r <= 3 ? 3 : r <= 5 ? 5 : r <= 8 ? 8 : 16; }

// Weights of band rows [y0, y0+rows) of one view: wb[ry][dy+r][x][PG], the
// NT words of a pixel's window row contiguous and padded to PG (a multiple
// of 4) so that a row is whole 16-byte words.  Block: 64 pixels x 4 window
// rows; each thread computes all taps of one (pixel, window row).  The
// window row's Lab pixels (columns x0-R-1 .. x0+64+R) are staged in LDS as
// packed words, and the block's output rows (each 64*PG contiguous words)
// are written back through LDS as whole 16-byte words by consecutive lanes.
// prox(dy, dx) = exp(-sqrt(dy^2+dx^2)/gp) comes from a per-block table; the
// colour term is exp2(|dLab| * kc) with kc = -log2(e)/gc (hardware exp2 and
// square root, each within an ulp or two: relative weight error ~1e-6).
template <bool HALF, int R>
__global__ void __launch_bounds__(256) k_aw_weights_band(const uint8_t *__restrict__ lab, uint32_t *__restrict__ wb,
                                                         int W, int H, int Wp, int y0, int nrow, int r, float kc,
                                                         float gp) {
    constexpr int NT = HALF ? R + 1 : 2 * R + 1, PG = (NT + 3) & ~3;
    constexpr int SW = 64 + 2 * R + 2;   // staged columns x0-R-1 .. x0+64+R
    constexpr uint32_t NONE = 0xFF000000u;   // staged column outside the image
    __shared__ float prox[4][2 * R + 2];
    __shared__ uint32_t pix[4][SW];
    __shared__ __attribute__((aligned(16))) uint32_t outw[4][64 * PG];
    const int e = 2 * r + 1;
    const int g = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.y * 4 + g;   // ry * e + (dy + r)
    const bool rowok = row < nrow;
    const int ry = row / e, dy = row - ry * e - r;
    const int y = y0 + ry, yy = y + dy;
    const int x0 = blockIdx.x * 64, x = x0 + lane;
    if (lane < 2 * R + 2) {
        const int dx = lane - R;
        prox[g][lane] = (dx > r || dx < -r || dy > r || dy < -r) ? 0.0f : expf(-sqrtf((float)(dy * dy + dx * dx)) / gp);
    }
    const bool yok = rowok && yy >= 0 && yy < H;
    for (int i = lane; i < SW; i += 64) {
        const int xx = x0 - R - 1 + i;
        uint32_t v = NONE;
        if (yok && xx >= 0 && xx < W) {
            const uint8_t *t = lab + 3 * ((size_t)yy * W + xx);
            v = (uint32_t)t[0] | (uint32_t)t[1] << 8 | (uint32_t)t[2] << 16;
        }
        pix[g][i] = v;
    }
    __syncthreads();
    const float *pr = prox[g];
    const bool ok = yok && x < W;
    uint32_t cpk = 0;   // the centre's Lab bytes
    if (ok) {
        const uint8_t *c = lab + 3 * ((size_t)y * W + x);
        cpk = (uint32_t)c[0] | (uint32_t)c[1] << 8 | (uint32_t)c[2] << 16;
    }
    const uint32_t cn = __builtin_amdgcn_udot4(cpk, cpk, 0u, false);   // |c|^2
    // The squared Lab distance |c - t|^2 = |c|^2 + |t|^2 - 2 c.t from byte
    // dot products: exact integers below 2^18, so the float square root
    // sees the value the float sum of the squared differences gives.
    auto weight = [&](int dx) {
        if (!ok || dx > R) return 0.0f;
        const uint32_t t = pix[g][lane + R + 1 + dx];
        if (t == NONE) return 0.0f;
        const int d2 = (int)__builtin_amdgcn_udot4(t, t, cn, false) - 2 * (int)__builtin_amdgcn_udot4(cpk, t, 0u, false);
        const float dist = __builtin_amdgcn_sqrtf((float)d2);
        return pr[dx + R] * __builtin_amdgcn_exp2f(dist * kc);
    };
    uint4 *mine = reinterpret_cast<uint4 *>(&outw[g][lane * PG]);
#pragma unroll
    for (int q = 0; q < PG / 4; q++) {
        uint32_t w4[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int jt = 4 * q + i;
            if (jt >= NT) {
                w4[i] = 0u;
            } else if (HALF) {
                h2 v;
                v.x = (_Float16)weight(2 * jt - R);
                v.y = (_Float16)weight(2 * jt - R + 1);
                w4[i] = __builtin_bit_cast(uint32_t, v);
            } else {
                w4[i] = __float_as_uint(weight(jt - R));
            }
        }
        mine[q] = make_uint4(w4[0], w4[1], w4[2], w4[3]);
    }
    __syncthreads();
    // row g of the block: 64*PG contiguous words at (row, x0)
    if (rowok) {
        const uint4 *src = reinterpret_cast<const uint4 *>(outw[g]);
        uint4 *dst = reinterpret_cast<uint4 *>(wb + ((size_t)row * Wp + x0) * PG);
        // (non-temporal: a band's weights exceed the caches and stream out)
        for (int i = lane; i < 16 * PG; i += 64)
            __builtin_nontemporal_store(__builtin_bit_cast(u32x4, src[i]), reinterpret_cast<u32x4 *>(dst) + i);
    }
}

// The value, hidden from constant folding (an empty asm that claims to
// modify it).  Keeps the compiler from re-deriving a per-access base
// address from a folded constant.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ int wrap_any(int v, int n) {
    if (v < 0) v += n;
    else if (v >= n) v -= n;
    return (v < 0 || v >= n) ? wrapn(v, n) : v;
}

// LDS layout of one aggregation tile: 64 pixels x 2 rows x 64 disparities,
// 8 waves.
template <bool HALF, int R>
struct AwTile {
    static constexpr int NTHR = 512;
    static constexpr int TXP = 64;                          // pixels per tile
    static constexpr int QW = TXP + 64;                     // target centres staged
    static constexpr int NT = HALF ? R + 1 : 2 * R + 1;     // weight words per (pixel, window row)
    static constexpr int PG = (NT + 3) & ~3;                // ... padded to 16-byte words
    static constexpr int CQ = (TXP + 2 * R + 2 + 3) / 4;    // cost quads per disparity row (<= 32)
    // Cost rows: FP32 floats; FP16 two copies of half pairs, one starting
    // at an even column and one at an odd column, so that every tap pair of
    // every pixel is one aligned word.  Odd pitch: no bank conflicts.
    static constexpr int PC = HALF ? (2 * CQ) | 1 : (4 * CQ) | 1;
    static constexpr int C_WORDS = ((HALF ? 2 : 1) * 64 * PC + 3) & ~3;
    static constexpr int T_WORDS = AW_NY * QW * PG;
    static constexpr int L_WORDS = AW_NY * TXP * PG;
    static constexpr int WORDS = (C_WORDS + T_WORDS + L_WORDS) > 64 * 65 ? C_WORDS + T_WORDS + L_WORDS : 64 * 65;
    static constexpr int SQ_C = 64 * 32 / NTHR;                        // 64 rows x 32 quad slots
    static constexpr int SQ_T = (AW_NY * QW * PG / 4 + NTHR - 1) / NTHR;
    static constexpr int SQ_L = (AW_NY * TXP * PG / 4 + NTHR - 1) / NTHR;
};

// Tile pixel of a wave's k-th lane column.  FP16: the two waves of a
// 16-pixel group take its even and its odd pixels, so that a lane's pixels
// share one column parity.  FP32: 8 consecutive pixels.
template <bool HALF>
__device__ __forceinline__ int aw_px(int wave, int k) {
    return HALF ? 16 * (wave >> 1) + 2 * k + (wave & 1) : 8 * wave + k;
}

template <bool HALF, int R>
__global__ void __launch_bounds__(512) k_aw_band(const uint32_t *__restrict__ wlb, const uint32_t *__restrict__ wtb,
                                                 const float *__restrict__ vol, float *__restrict__ out, int W, int H,
                                                 int D, int Wp, int y0, int rows, int r, int nxb, int nyb, int ndb) {
    using T = AwTile<HALF, R>;
    constexpr int NT = T::NT, PG = T::PG, NTHR = T::NTHR, TXP = T::TXP, QW = T::QW;
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];
    uint32_t *Ct = lds, *Tt = lds + T::C_WORDS, *Lt = lds + T::C_WORDS + T::T_WORDS;

    const int nblk = nxb * nyb * ndb;
    const int lin = xcd_order(blockIdx.x, nblk);
    const int xb = lin % nxb, rest = lin / nxb;
    const int db = rest % ndb, yb = rest / ndb;
    const int x0 = xb * TXP, d0 = db * 64, ry0 = yb * AW_NY, yt = y0 + ry0;
    const int tid = threadIdx.x, dl = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nrows = min(AW_NY, rows - ry0);
    const int e = 2 * r + 1;
    const size_t pl = plane(W, H);
    const int q_lo = x0 - d0 - 64;   // first target centre staged
    const bool fast_c = (W & 3) == 0 && ((x0 - R) & 3) == 0 && x0 - R >= 0 && x0 - R + 4 * T::CQ <= W;

    auto row_valid = [&](int yy, int ny) {
        const int dy = yy - (yt + ny);
        return ny < nrows && dy >= -r && dy <= r;
    };
    // Weight row (pixel or centre c) of tile row ny for window row yy.
    auto wrow = [&](const uint32_t *base, int yy, int ny, int c) {
        return base + (((size_t)(ry0 + ny) * e + (yy - (yt + ny) + r)) * Wp + c) * PG;
    };

    // Global -> registers -> LDS for window row yy: the cost tile (columns
    // x0-R .. of the block's 64 disparities, wrapped) and the target weight
    // rows of the QW centres from x0-d0-64 (wrapped); FP32 also the
    // reference weight rows of the tile's pixels.  The loads for the next
    // row are issued before the arithmetic on this one.
    float4 rc[T::SQ_C];
    uint4 rt[T::SQ_T], rl[T::SQ_L];
    auto load = [&](int yy) {
        const float *crow = vol + (size_t)wrap_any(yy, H) * W;
#pragma unroll
        for (int s = 0; s < T::SQ_C; s++) {
            const int u = s * NTHR + tid, dd = u >> 5, qc = u & 31, d = d0 + dd;
            float4 v = {0.0f, 0.0f, 0.0f, 0.0f};
            if (qc < T::CQ && d < D) {
                const float *src = crow + (size_t)d * pl;
                const int c0 = x0 - R + 4 * qc;
                if (fast_c) {
                    v = *reinterpret_cast<const float4 *>(src + c0);
                } else {
                    v.x = src[wrap_any(c0, W)];
                    v.y = src[wrap_any(c0 + 1, W)];
                    v.z = src[wrap_any(c0 + 2, W)];
                    v.w = src[wrap_any(c0 + 3, W)];
                }
            }
            rc[s] = v;
        }
#pragma unroll
        for (int s = 0; s < T::SQ_T; s++) {
            const int u = s * NTHR + tid, ny = u / (QW * PG / 4), v4 = u % (QW * PG / 4);
            const int q = v4 / (PG / 4), g = v4 % (PG / 4);
            uint4 v = {0u, 0u, 0u, 0u};
            if (u < AW_NY * QW * PG / 4 && row_valid(yy, ny))
                v = reinterpret_cast<const uint4 *>(wrow(wtb, yy, ny, wrap_any(q_lo + q, W)))[g];
            rt[s] = v;
        }
#pragma unroll
        for (int s = 0; s < T::SQ_L; s++) {
            const int u = s * NTHR + tid, ny = u / (TXP * PG / 4), v4 = u % (TXP * PG / 4);
            const int px = v4 / (PG / 4), g = v4 % (PG / 4);
            uint4 v = {0u, 0u, 0u, 0u};
            if (u < AW_NY * TXP * PG / 4 && row_valid(yy, ny))
                v = reinterpret_cast<const uint4 *>(wrow(wlb, yy, ny, x0 + px))[g];
            rl[s] = v;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int s = 0; s < T::SQ_C; s++) {
            const int u = s * NTHR + tid, dd = u >> 5, qc = u & 31;
            if (qc >= T::CQ) continue;
            const float4 v = rc[s];
            if (HALF) {
                _Float16 *even = reinterpret_cast<_Float16 *>(Ct + dd * T::PC);
                _Float16 *odd = reinterpret_cast<_Float16 *>(Ct + (64 + dd) * T::PC) - 1;   // half i holds column i+1
                h2 a, b;
                a.x = (_Float16)v.x;
                a.y = (_Float16)v.y;
                b.x = (_Float16)v.z;
                b.y = (_Float16)v.w;
                reinterpret_cast<h2 *>(even)[2 * qc] = a;
                reinterpret_cast<h2 *>(even)[2 * qc + 1] = b;
                if (qc > 0) odd[4 * qc] = a.x;
                odd[4 * qc + 1] = a.y;
                odd[4 * qc + 2] = b.x;
                odd[4 * qc + 3] = b.y;
            } else {
                uint32_t *dst = Ct + dd * T::PC + 4 * qc;
                dst[0] = __float_as_uint(v.x);
                dst[1] = __float_as_uint(v.y);
                dst[2] = __float_as_uint(v.z);
                dst[3] = __float_as_uint(v.w);
            }
        }
#pragma unroll
        for (int s = 0; s < T::SQ_T; s++) {
            const int u = s * NTHR + tid;
            if (u < AW_NY * QW * PG / 4) reinterpret_cast<uint4 *>(Tt)[u] = rt[s];
        }
#pragma unroll
        for (int s = 0; s < T::SQ_L; s++) {
            const int u = s * NTHR + tid;
            if (u < AW_NY * TXP * PG / 4) reinterpret_cast<uint4 *>(Lt)[u] = rl[s];
        }
    };

    float num[AW_NY][AW_NX], den[AW_NY][AW_NX];
#pragma unroll
    for (int a = 0; a < AW_NY; a++)
#pragma unroll
        for (int k = 0; k < AW_NX; k++) num[a][k] = den[a][k] = 0.0f;

    // FP16: per (lane column k, tile row ny) both weight rows come from LDS
    // as 16-byte reads (the reference row is the same address for all lanes
    // of the wave); all products are formed before the dot products, since
    // a dot2 that reads the product written just before it costs a wait
    // state.
    auto compute_half = [&](int yy, const h2 *cw) {
        const h2 one = {(_Float16)1.0f, (_Float16)1.0f};
#pragma unroll
        for (int ny = 0; ny < AW_NY; ny++) {
            if (!row_valid(yy, ny)) continue;
#pragma unroll
            for (int k = 0; k < AW_NX; k++) {
                const int px = aw_px<true>(wave, k);
                const uint4 *tw =
                    reinterpret_cast<const uint4 *>(lds + opaque(T::C_WORDS + (ny * QW + px - dl + 64) * PG));
                const uint4 *lw = reinterpret_cast<const uint4 *>(Lt + (ny * TXP + px) * PG);
                uint4 tv[PG / 4], lv[PG / 4];
#pragma unroll
                for (int g = 0; g < PG / 4; g++) {
                    tv[g] = tw[g];
                    lv[g] = lw[g];
                }
                // The target row's last 16-byte word is read whole even where
                // only its first words are taps: the lanes' rows lie PG words
                // apart, and 4-byte reads at that stride share banks, where
                // 16-byte reads do not.
                if constexpr (NT % 4 != 0) asm volatile("" ::"v"(__builtin_bit_cast(u32x4, tv[PG / 4 - 1])));
                const uint32_t *t32 = reinterpret_cast<const uint32_t *>(tv);
                const uint32_t *l32 = reinterpret_cast<const uint32_t *>(lv);
                h2 p[NT];
#pragma unroll
                for (int j = 0; j < NT; j++) p[j] = __builtin_bit_cast(h2, l32[j]) * __builtin_bit_cast(h2, t32[j]);
                float sn = num[ny][k], sd = den[ny][k];
#pragma unroll
                for (int j = 0; j < NT; j++) {
                    sn = __builtin_amdgcn_fdot2(p[j], cw[k + j], sn, false);
                    sd = __builtin_amdgcn_fdot2(p[j], one, sd, false);
                }
                num[ny][k] = sn;
                den[ny][k] = sd;
            }
        }
    };

    auto compute = [&](int yy) {
        const bool v0 = row_valid(yy, 0), v1 = row_valid(yy, 1);
        if (!v0 && !v1) return;
        if (HALF) {
            // Tap pair j of lane column k covers tile columns px+2j, px+2j+1;
            // px = 16*(wave>>1) + parity + 2k, so it is word 8*(wave>>1)+k+j
            // of the parity's copy.
            h2 cw[AW_NX + R];
            const uint32_t *crow = Ct + ((wave & 1) * 64 + dl) * T::PC + 8 * (wave >> 1);
#pragma unroll
            for (int i = 0; i < AW_NX + R; i++) cw[i] = __builtin_bit_cast(h2, crow[i]);
            compute_half(yy, cw);
        } else {
            float cf[AW_NX + 2 * R];
            const float *crow = reinterpret_cast<const float *>(Ct) + dl * T::PC + 8 * wave;
#pragma unroll
            for (int i = 0; i < AW_NX + 2 * R; i++) cf[i] = crow[i];
#pragma unroll
            for (int ny = 0; ny < AW_NY; ny++) {
                if (!row_valid(yy, ny)) continue;
#pragma unroll
                for (int k = 0; k < AW_NX; k++) {
                    const int px = aw_px<false>(wave, k);
                    const float4 *tw =
                        reinterpret_cast<const float4 *>(lds + opaque(T::C_WORDS + (ny * QW + px - dl + 64) * PG));
                    const float4 *lw = reinterpret_cast<const float4 *>(Lt + (ny * TXP + px) * PG);
                    float4 tv[PG / 4], lv[PG / 4];
#pragma unroll
                    for (int g = 0; g < PG / 4; g++) {
                        tv[g] = tw[g];
                        lv[g] = lw[g];
                    }
                    // the last target word read whole (as in the FP16 form)
                    if constexpr (NT % 4 != 0) asm volatile("" ::"v"(__builtin_bit_cast(u32x4, tv[PG / 4 - 1])));
                    const float *tf = reinterpret_cast<const float *>(tv);
                    const float *lf = reinterpret_cast<const float *>(lv);
                    float sn = num[ny][k], sd = den[ny][k];
#pragma unroll
                    for (int t = 0; t < NT; t++) {
                        const float p = lf[t] * tf[t];
                        sd += p;
                        sn = __builtin_fmaf(p, cf[k + t], sn);
                    }
                    num[ny][k] = sn;
                    den[ny][k] = sd;
                }
            }
        }
    };

    // FP16 prefetches the next window row into registers during the
    // arithmetic; FP32 (more live registers) loads and stores back to back.
    const int ylast = yt + nrows - 1 + r;
    if (HALF) load(yt - r);
    for (int yy = yt - r; yy <= ylast; yy++) {
        __syncthreads();
        if (!HALF) load(yy);
        store();
        __syncthreads();
        if (HALF && yy < ylast) load(yy + 1);
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
        for (int k = 0; k < AW_NX; k++) tile[dl * 65 + aw_px<HALF>(wave, k)] = num[ny][k] / den[ny][k];
        __syncthreads();
        const int y = yt + ny;
        for (int i = tid; i < 64 * TXP; i += NTHR) {
            const int dd = i / TXP, px = i % TXP;
            const int d = d0 + dd, x = x0 + px;
            if (d < D && x < W) stream_store(&out[(size_t)d * pl + (size_t)y * W + x], tile[dd * 65 + px]);
        }
    }
}

// FP16 aggregation with its staging by direct loads into LDS (global_load_lds:
// no registers, no LDS store instructions), double-buffered so that window
// row yy+1 arrives while row yy is summed.  A block is one tile row of 64
// pixels x 64 disparities (lane = disparity, 8 waves; two blocks share a CU);
// the numerators use the arithmetic of k_aw_band<true, R> (half products,
// dot2 into float), the denominators run on the matrix cores.  The cost rows
// stay FP32 in LDS (a pitch at which the per-lane row reads meet no bank
// conflicts) and each lane packs its register window of half pairs with
// v_cvt_pk_f16_f32, which rounds as the staged conversion did.
template <int R, int NY, bool C4>
struct AwDma {
    static constexpr int NWAVE = 8, NTHR = 64 * NWAVE;
    static constexpr int TXP = 64, QW = TXP + 64;
    static constexpr int NT = R + 1, PG = (NT + 3) & ~3;
    // FP32 cost columns per disparity row: the floats a lane's pairs reach
    // (up to 16*3 + 1 + 2*(AW_NX+R) - 1).  C4: rows of whole 16-byte words,
    // loaded as such, at a pitch that is an odd multiple of 4 words (16-byte
    // reads of the 64 lanes' rows then meet no bank conflicts); otherwise
    // single words at an odd pitch.
    static constexpr int NCF = 49 + 2 * (AW_NX + R);
    static constexpr int PCF = C4 ? (((NCF + 3) / 4) | 1) * 4 : NCF | 1;
    static constexpr int C_INS = C4 ? 64 * PCF / 256 : PCF;             // loads of 64 lanes
    static constexpr int T_QUADS = NY * QW * PG / 4, L_QUADS = NY * TXP * PG / 4;
    static constexpr int T_INS = (T_QUADS + 63) / 64, L_INS = (L_QUADS + 63) / 64;   // 16-byte loads
    static constexpr int C_PW = (C_INS + NWAVE - 1) / NWAVE, T_PW = (T_INS + NWAVE - 1) / NWAVE,
                         L_PW = (L_INS + NWAVE - 1) / NWAVE;
    static constexpr int T_OFF = 64 * PCF, L_OFF = T_OFF + T_INS * 256;
    // (each buffer also holds, at the end, the 64 x 129 denominator matrix
    // or the 64 x 65 output stage)
    static constexpr int WORDS = L_OFF + L_INS * 256 > 64 * 129 ? L_OFF + L_INS * 256 : 64 * 129;
    static_assert(2 * WORDS * 4 <= 160 * 1024, "two buffers must fit the LDS");
};

// Direct loads into LDS (64 lanes: lane i's 4 or 16 bytes from base + its
// byte offset land at LDS byte address lds + 4i or 16i), written out so that
// the compiler neither tracks nor waits for them: the kernel issues its own
// vmcnt waits.  M0 holds the LDS address; nothing else in these kernels
// uses M0.
typedef __attribute__((address_space(3))) uint32_t lds_u32;
__device__ __forceinline__ uint32_t lds_addr(const uint32_t *p) { return (uint32_t)(size_t)(const lds_u32 *)p; }
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
__device__ __forceinline__ void dma_b32(const void *base, uint32_t off, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dword %0, %1" ::"v"(off), "s"(base), "s"(lds) : "memory", "m0");
}
__device__ __forceinline__ void dma_b128(const void *base, uint32_t off, uint32_t lds) {
    asm volatile("s_mov_b32 m0, %2\n\tglobal_load_lds_dwordx4 %0, %1" ::"v"(off), "s"(base), "s"(lds)
                 : "memory", "m0");
}
#pragma clang diagnostic pop

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
    // s_waitcnt with vmcnt N (bits 3:0 and 15:14), expcnt and lgkmcnt at
    // their maxima (no wait on them)
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// (four waves per SIMD: two blocks per CU, so at most 128 registers)
template <int R, int NY, bool C4>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) k_aw_band_dma(const uint32_t *__restrict__ wlb, const uint32_t *__restrict__ wtb,
                                                     const float *__restrict__ vol, float *__restrict__ out, int W,
                                                     int H, int D, int Wp, int y0, int rows, int r, int nxb, int nyb,
                                                     int ndb) {
    using T = AwDma<R, NY, C4>;
    static_assert(NY == 1, "the matrix-core denominators cover one tile row");
    constexpr int NT = T::NT, PG = T::PG, TXP = T::TXP, QW = T::QW, PCF = T::PCF;
    __shared__ __attribute__((aligned(16))) uint32_t lbuf[2 * T::WORDS];
    uint32_t *const bufA = lbuf, *const bufB = lbuf + T::WORDS;

    const int nblk = nxb * nyb * ndb;
    const int lin = xcd_order(blockIdx.x, nblk);
    const int xb = lin % nxb, rest = lin / nxb;
    const int db = rest % ndb, yb = rest / ndb;
    const int x0 = xb * TXP, d0 = db * 64, ry0 = yb * NY, yt = y0 + ry0;
    const int tid = threadIdx.x, dl = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nrows = min(NY, rows - ry0);
    const int e = 2 * r + 1;
    const size_t pl = plane(W, H);
    const int q_lo = x0 - d0 - 64;   // first target centre staged

    // workspace row of tile row ny for window row yy, clamped into the
    // tile row's own rows when yy lies outside its window (data then unused)
    auto wrow_index = [&](int yy, int ny) {
        const int nyc = min(ny, nrows - 1);
        const int dy = min(max(yy - (yt + nyc), -r), r);
        return (size_t)(ry0 + nyc) * e + (dy + r);
    };

    // Per-lane byte offsets of the loads, fixed for the block.
    uint32_t offc[T::C_PW], offt[T::T_PW], offl[T::L_PW];
    uint32_t nyt = 0, nyl = 0;   // bit k: the k-th weight load belongs to tile row 1
#pragma unroll
    for (int k = 0; k < T::C_PW; k++) {
        const int ins = min(wave + T::NWAVE * k, T::C_INS - 1);
        const int w = (ins * 64 + dl) * (C4 ? 4 : 1);   // first staged word of the lane
        const int dd = min(w / PCF, 63), j = w - (w / PCF) * PCF;
        const int d = min(d0 + dd, D - 1);
        offc[k] = 4u * ((uint32_t)((size_t)(d - d0) * pl) + (uint32_t)wrap_any(x0 - R + j, W));
    }
#pragma unroll
    for (int k = 0; k < T::T_PW; k++) {
        const int ins = min(wave + T::NWAVE * k, T::T_INS - 1);
        const int u = min(ins * 64 + dl, T::T_QUADS - 1);
        const int ny = u / (QW * PG / 4), v4 = u - ny * (QW * PG / 4);
        const int q = v4 / (PG / 4), g = v4 - q * (PG / 4);
        offt[k] = 4u * ((uint32_t)wrap_any(q_lo + q, W) * PG + 4 * g);
        nyt |= (uint32_t)ny << k;
    }
#pragma unroll
    for (int k = 0; k < T::L_PW; k++) {
        const int ins = min(wave + T::NWAVE * k, T::L_INS - 1);
        const int u = min(ins * 64 + dl, T::L_QUADS - 1);
        const int ny = u / (TXP * PG / 4), v4 = u - ny * (TXP * PG / 4);
        const int px = v4 / (PG / 4), g = v4 - px * (PG / 4);
        offl[k] = 4u * ((uint32_t)(x0 + px) * PG + 4 * g);
        nyl |= (uint32_t)ny << k;
    }

    // Window row yy into buffer BUF: the FP32 cost rows of the block's 64
    // disparities (columns x0-R .., wrapped), the target weight rows of the
    // QW centres from x0-d0-64 and the reference rows of the tile's pixels.
    // The weight rows of tile row 1 start (row 1's index - row 0's) rows of
    // Wp*PG words after row 0's; that distance is added to the lanes whose
    // load belongs to tile row 1 (uniform per window row).
    auto issue = [&](int which, int yy) {
        const uint32_t *buf = lbuf + which * T::WORDS;
        const float *crow = vol + (size_t)d0 * pl + (size_t)wrap_any(yy, H) * W;
#pragma unroll
        for (int k = 0; k < T::C_PW; k++) {
            const int ins = min(wave + T::NWAVE * k, T::C_INS - 1);
            if constexpr (C4)
                dma_b128(crow, offc[k], lds_addr(buf + ins * 256));
            else
                dma_b32(crow, offc[k], lds_addr(buf + ins * 64));
        }
        const size_t i0 = wrow_index(yy, 0), i1 = wrow_index(yy, 1);
        const uint32_t step1 = (uint32_t)((i1 - i0) * (size_t)Wp * PG * 4);
        const uint32_t *tb0 = wtb + i0 * Wp * PG, *lb0 = wlb + i0 * Wp * PG;
#pragma unroll
        for (int k = 0; k < T::T_PW; k++) {
            const int ins = min(wave + T::NWAVE * k, T::T_INS - 1);
            dma_b128(tb0, offt[k] + ((nyt >> k) & 1 ? step1 : 0u), lds_addr(buf + T::T_OFF + ins * 256));
        }
#pragma unroll
        for (int k = 0; k < T::L_PW; k++) {
            const int ins = min(wave + T::NWAVE * k, T::L_INS - 1);
            dma_b128(lb0, offl[k] + ((nyl >> k) & 1 ? step1 : 0u), lds_addr(buf + T::L_OFF + ins * 256));
        }
    };

    // Numerators on the vector units, one per lane column; the
    // denominators on the matrix cores.  For the tile row, den(px, d) is
    // entry (px, c = px - d + 64) of the product of the reference rows
    // (64 pixels x 2*PG taps, half) and the transposed target rows (128
    // centres): summed over window rows, a 64 x 128 float matrix whose
    // band c - px in [1, 64] holds every denominator of the tile.  Waves
    // 0-5 each keep one 32 x 32 block of the band's six in accumulator
    // registers (v_mfma_f32_32x32x16_f16 and 32x32x8f16: products exact,
    // float sums); the zero taps that pad a row to PG words add nothing.
    float num[AW_NX];
#pragma unroll
    for (int k = 0; k < AW_NX; k++) num[k] = 0.0f;
    const bool has_tile = wave < 6;
    const int mt = wave / 3, nt = mt + wave % 3;   // pixel block, centre block
    f32x16 macc = {};

    // The sums of window row yy from buffer BUF.  Tap pair j of lane
    // column k covers tile columns px+2j, px+2j+1 with px = 16*(wave>>1) +
    // parity + 2k: floats 16*(wave>>1) + parity + 2(k+j) and the next of the
    // lane's staged row.
    auto compute = [&](int which) {
        const uint32_t *buf = lbuf + which * T::WORDS;
        h2 cw[AW_NX + R];
        if constexpr (C4) {
            // the lane's floats from 16*(wave>>1) as whole 16-byte words,
            // then the pairs from the parity's offset
            constexpr int NQ = (2 * (AW_NX + R) + 1 + 3) / 4;
            const float4 *cq = reinterpret_cast<const float4 *>(buf + dl * PCF + 16 * (wave >> 1));
            float cf[4 * NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                const float4 v = cq[q];
                cf[4 * q] = v.x;
                cf[4 * q + 1] = v.y;
                cf[4 * q + 2] = v.z;
                cf[4 * q + 3] = v.w;
            }
            const bool par = wave & 1;
#pragma unroll
            for (int i = 0; i < AW_NX + R; i++) {
                h2 v;
                v.x = (_Float16)(par ? cf[2 * i + 1] : cf[2 * i]);
                v.y = (_Float16)(par ? cf[2 * i + 2] : cf[2 * i + 1]);
                cw[i] = v;
            }
        } else {
            const float *cr = reinterpret_cast<const float *>(buf) + dl * PCF + 16 * (wave >> 1) + (wave & 1);
#pragma unroll
            for (int i = 0; i < AW_NX + R; i++) {
                h2 v;
                v.x = (_Float16)cr[2 * i];
                v.y = (_Float16)cr[2 * i + 1];
                cw[i] = v;
            }
        }
        const uint32_t *Lt = buf + T::L_OFF, *Tt = buf + T::T_OFF;
        if (has_tile) {
            // lane l: pixel row / centre column l % 32; taps in groups of 16
            // (v_mfma_f32_32x32x16_f16, lane l: taps 8*(l/32) .. +7, whole
            // 16-byte words at the rows' stride, no bank conflicts), then a
            // last group of 8 where a row has 8 more (32x32x8, lane l: taps
            // 4*(l/32) .. +3)
            const uint32_t *ar = Lt + (32 * mt + (dl & 31)) * PG, *br = Tt + (32 * nt + (dl & 31)) * PG;
#pragma unroll
            for (int ks = 0; ks < PG / 8; ks++) {
                const half8 av = *reinterpret_cast<const half8 *>(ar + 8 * ks + 4 * (dl >> 5));
                const half8 bv = *reinterpret_cast<const half8 *>(br + 8 * ks + 4 * (dl >> 5));
                macc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, macc, 0, 0, 0);
            }
            if constexpr (PG % 8 != 0) {
                const half4 av = *reinterpret_cast<const half4 *>(ar + PG - 4 + 2 * (dl >> 5));
                const half4 bv = *reinterpret_cast<const half4 *>(br + PG - 4 + 2 * (dl >> 5));
                macc = __builtin_amdgcn_mfma_f32_32x32x8f16(av, bv, macc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int k = 0; k < AW_NX; k++) {
            const int px = aw_px<true>(wave, k);
            const uint4 *tw = reinterpret_cast<const uint4 *>(buf + opaque(T::T_OFF + (px - dl + 64) * PG));
            const uint4 *lw = reinterpret_cast<const uint4 *>(Lt + px * PG);
            uint4 tv[PG / 4], lv[PG / 4];
#pragma unroll
            for (int g = 0; g < PG / 4; g++) {
                tv[g] = tw[g];
                lv[g] = lw[g];
            }
            // The target row's last 16-byte word is read whole even where
            // only its first words are taps: the lanes' rows lie PG words
            // apart, and 4-byte reads at that stride share banks, where
            // 16-byte reads do not.
            if constexpr (NT % 4 != 0) asm volatile("" ::"v"(__builtin_bit_cast(u32x4, tv[PG / 4 - 1])));
            const uint32_t *t32 = reinterpret_cast<const uint32_t *>(tv);
            const uint32_t *l32 = reinterpret_cast<const uint32_t *>(lv);
            h2 p[NT];
#pragma unroll
            for (int j = 0; j < NT; j++) p[j] = __builtin_bit_cast(h2, l32[j]) * __builtin_bit_cast(h2, t32[j]);
            float sn = num[k];
#pragma unroll
            for (int j = 0; j < NT; j++) sn = __builtin_amdgcn_fdot2(p[j], cw[k + j], sn, false);
            num[k] = sn;
        }
    };

    const int ylast = yt + nrows - 1 + r;
    // Each step: every wave's reads of the previous row have returned
    // (lgkmcnt 0) and its loads of this row have landed (vmcnt 0); past the
    // barrier the row is complete in the buffer and the other buffer is free
    // for the next row's loads, which fly during this row's sums.
    auto step = [&](int cur, int yy) {
        __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
        wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        if (yy < ylast) issue(cur ^ 1, yy + 1);
        compute(cur);
    };
    issue(0, yt - r);
    for (int yy = yt - r; yy <= ylast; yy++) step((yy - yt + r) & 1, yy);

    // The denominators: each wave's accumulator block into a 64 x 129
    // float matrix (odd pitch: a lane's band entry px*129 + px - d + 64 is
    // one word from its neighbour's), then each lane takes its pixels'.
    float *dm = reinterpret_cast<float *>(bufB);
    __syncthreads();
    if (has_tile) {
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int i = 32 * mt + 8 * (q / 4) + 4 * (dl >> 5) + (q % 4), c = 32 * nt + (dl & 31);
            dm[i * 129 + c] = macc[q];
        }
    }
    __syncthreads();
    // Write through LDS so that each disparity row of 64 pixels is stored
    // contiguously.
    float *tile = reinterpret_cast<float *>(bufA);
#pragma unroll
    for (int k = 0; k < AW_NX; k++) {
        const int px = aw_px<true>(wave, k);
        tile[dl * 65 + px] = num[k] / dm[px * 129 + px - dl + 64];
    }
    __syncthreads();
    for (int i = tid; i < 64 * TXP; i += T::NTHR) {
        const int dd = i / TXP, px = i % TXP;
        const int d = d0 + dd, x = x0 + px;
        if (d < D && x < W) stream_store(&out[(size_t)d * pl + (size_t)yt * W + x], tile[dd * 65 + px]);
    }
}

// FP16 aggregation from an FP16 cost volume, M disparities per lane.  A block
// is one tile row of 64 pixels x 64*M disparities (8 waves).  Lane l owns
// disparities b, b+2, .., b+2(M-1) with b = 2M*(l/2) + l%2, and a wave walks
// the columns px, px+2, .., so the target centre of (px, b+2j) is the one of
// (px-2j, b): each column brings one new target weight row per lane, the
// other M-1 come from the previous columns' registers, and the column's
// reference row serves all M outputs.  Against one disparity per lane this
// reads 1/M of the weight bytes from LDS per output.  The cost rows come
// from the FP16 volume (whole 16-byte words, staged from column x0 - R8);
// each lane forms its tap pairs with one byte permute per pair.  Products,
// sums, their order and the matrix-core denominators are those of
// k_aw_band_dma, so the two kernels' outputs are bit-identical when the
// FP16 volume holds the FP32 costs rounded to nearest.
template <int R, int M>
struct AwHalf {
    static constexpr int NWAVE = 8, NTHR = 64 * NWAVE;
    static constexpr int TXP = 64, DT = 64 * M, QW = TXP + DT;
    static constexpr int NT = R + 1, PG = (NT + 3) & ~3;
    static constexpr int R8 = (R + 7) & ~7;   // the staged row starts at x0 - R8 (16-byte aligned)
    // halves of a cost row that the lanes reach: from x0 - R8 up to column
    // 16*3 + 1 + 2*(AW_NX-1) + R + 1 of the tile
    static constexpr int NCH = R8 - R + 49 + 2 * (AW_NX + R);
    static constexpr int PCQ = ((NCH + 7) / 8) | 1;   // row pitch in 16-byte words, odd
    static constexpr int PCW = 4 * PCQ;
    // 16-byte words a lane reads of one row (from word 2*(wave/2))
    static constexpr int NQ = (R8 - R + 1 + 2 * (AW_NX + R) + 7) / 8;
    static constexpr int C_INS = DT * PCQ / 64;          // 16-byte loads of 64 lanes
    static constexpr int T_INS = QW * PG / 4 / 64, L_INS = TXP * PG / 4 / 64;
    static constexpr int C_PW = (C_INS + NWAVE - 1) / NWAVE, T_PW = (T_INS + NWAVE - 1) / NWAVE,
                         L_PW = (L_INS + NWAVE - 1) / NWAVE;
    static constexpr int T_OFF = DT * PCW, L_OFF = T_OFF + QW * PG, WORDS = L_OFF + TXP * PG;
    static constexpr int DMP = QW + 1;                   // pitch of the denominator matrix (odd)
    static constexpr int NBLK = 2 * (2 * M + 1), NSLOT = (NBLK + NWAVE - 1) / NWAVE;
    static constexpr int LBUF_A = 2 * WORDS > 64 * DMP ? 2 * WORDS : 64 * DMP;
    static constexpr int LBUF = LBUF_A > DT * 65 ? LBUF_A : DT * 65;
    static_assert(QW * PG % 256 == 0 && TXP * PG % 256 == 0 && DT * PCQ % 64 == 0, "whole loads");
    static_assert(6 + NQ <= PCQ, "a lane's words lie in its row");
    static_assert(((R8 - R) >> 1) + AW_NX + R < 4 * NQ, "a lane's pairs lie in its words");
    static_assert(LBUF * 4 <= 160 * 1024, "the buffers must fit the LDS");
};

template <int R, int M>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(M == 1 ? 4 : 2)))
k_aw_band_h(const uint32_t *__restrict__ wlb, const uint32_t *__restrict__ wtb, const uint16_t *__restrict__ volh,
            float *__restrict__ out, int W, int H, int D, int Wp, int y0, int r, int nxb, int nyb, int ndb) {
    using T = AwHalf<R, M>;
    constexpr int NT = T::NT, PG = T::PG, DT = T::DT, PCW = T::PCW, PCQ = T::PCQ, DMP = T::DMP;
    __shared__ __attribute__((aligned(16))) uint32_t lbuf[T::LBUF];

    const int nblk = nxb * nyb * ndb;
    const int lin = xcd_order(blockIdx.x, nblk);
    const int xb = lin % nxb, rest = lin / nxb;
    const int db = rest % ndb, yb = rest / ndb;
    const int x0 = xb * 64, d0 = db * DT, yt = y0 + yb;
    const int tid = threadIdx.x, dl = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int e = 2 * r + 1;
    const size_t pl = plane(W, H);
    const int q_lo = x0 - d0 - DT;   // target centre of LDS row 0

    // Per-lane byte offsets of the loads, fixed for the block.
    uint32_t offc[T::C_PW], offt[T::T_PW], offl[T::L_PW];
#pragma unroll
    for (int k = 0; k < T::C_PW; k++) {
        const int u = min(wave + T::NWAVE * k, T::C_INS - 1) * 64 + dl;   // 16-byte word of the tile's rows
        const int dd = u / PCQ, j = u - dd * PCQ;
        const int d = min(d0 + dd, D - 1);
        offc[k] = 2u * ((uint32_t)((size_t)(d - d0) * pl) + (uint32_t)wrap_any(x0 - T::R8 + 8 * j, W));
    }
#pragma unroll
    for (int k = 0; k < T::T_PW; k++) {
        const int u = min(wave + T::NWAVE * k, T::T_INS - 1) * 64 + dl;
        const int q = u / (PG / 4), g = u - q * (PG / 4);
        offt[k] = 4u * ((uint32_t)wrap_any(q_lo + q, W) * PG + 4 * g);
    }
#pragma unroll
    for (int k = 0; k < T::L_PW; k++) {
        const int u = min(wave + T::NWAVE * k, T::L_INS - 1) * 64 + dl;
        const int px = u / (PG / 4), g = u - px * (PG / 4);
        offl[k] = 4u * ((uint32_t)(x0 + px) * PG + 4 * g);
    }

    // Window row yy into buffer WHICH: the FP16 cost rows of the block's DT
    // disparities (columns x0-R8 .., wrapped), the target weight rows of the
    // QW centres from q_lo and the reference rows of the tile's pixels.
    auto issue = [&](int which, int yy) {
        const uint32_t *buf = lbuf + which * T::WORDS;
        const uint16_t *crow = volh + (size_t)d0 * pl + (size_t)wrap_any(yy, H) * W;
#pragma unroll
        for (int k = 0; k < T::C_PW; k++)
            dma_b128(crow, offc[k], lds_addr(buf + min(wave + T::NWAVE * k, T::C_INS - 1) * 256));
        const size_t wr = (size_t)yb * e + (yy - yt + r);
        const uint32_t *tb = wtb + wr * Wp * PG, *lb = wlb + wr * Wp * PG;
#pragma unroll
        for (int k = 0; k < T::T_PW; k++)
            dma_b128(tb, offt[k], lds_addr(buf + T::T_OFF + min(wave + T::NWAVE * k, T::T_INS - 1) * 256));
#pragma unroll
        for (int k = 0; k < T::L_PW; k++)
            dma_b128(lb, offl[k], lds_addr(buf + T::L_OFF + min(wave + T::NWAVE * k, T::L_INS - 1) * 256));
    };

    const int base = 2 * M * (dl >> 1) + (dl & 1);   // the lane's first disparity in the tile
    const int pxw = 16 * (wave >> 1) + (wave & 1);    // the wave's first column
    // Tap pair i of column k and disparity row j: halves
    // R8 - R + (wave & 1) + 2(k + i) and the next of the lane's words from
    // 16-byte word 2*(wave/2) of the row; f halves past word R8-R>>1 + k + i.
    constexpr int CB = (T::R8 - R) >> 1;
    const int f = ((T::R8 - R) & 1) + (wave & 1);
    const uint32_t psel = f == 0 ? 0x03020100u : f == 1 ? 0x05040302u : 0x07060504u;
    float num[AW_NX][M];
#pragma unroll
    for (int k = 0; k < AW_NX; k++)
#pragma unroll
        for (int j = 0; j < M; j++) num[k][j] = 0.0f;
    f32x16 macc[T::NSLOT];
#pragma unroll
    for (int s = 0; s < T::NSLOT; s++) macc[s] = f32x16{};

    auto compute = [&](int which, int yy) {
        const uint32_t *buf = lbuf + which * T::WORDS;
        uint32_t cw[M][AW_NX + R];
#pragma unroll
        for (int j = 0; j < M; j++) {
            const u32x4 *cq = reinterpret_cast<const u32x4 *>(buf + (base + 2 * j) * PCW + 8 * (wave >> 1));
            uint32_t wd[4 * T::NQ];
#pragma unroll
            for (int q = 0; q < T::NQ; q++) {
                const u32x4 v = cq[q];
                wd[4 * q] = v.x;
                wd[4 * q + 1] = v.y;
                wd[4 * q + 2] = v.z;
                wd[4 * q + 3] = v.w;
            }
#pragma unroll
            for (int i = 0; i < AW_NX + R; i++)
                cw[j][i] = __builtin_amdgcn_perm(wd[CB + i + 1], wd[CB + i], psel);
        }
        const uint32_t *Lt = buf + T::L_OFF, *Tt = buf + T::T_OFF;
        // denominators: block b = wave + 8s of the band's NBLK, pixel block
        // mt, centre block nt (as in k_aw_band_dma)
#pragma unroll
        for (int s = 0; s < T::NSLOT; s++) {
            const int b = wave + T::NWAVE * s;
            if (b < T::NBLK) {
                const int mt = b / (2 * M + 1), nt = mt + b % (2 * M + 1);
                const uint32_t *ar = Lt + (32 * mt + (dl & 31)) * PG, *br = Tt + (32 * nt + (dl & 31)) * PG;
#pragma unroll
                for (int ks = 0; ks < PG / 8; ks++) {
                    const half8 av = *reinterpret_cast<const half8 *>(ar + 8 * ks + 4 * (dl >> 5));
                    const half8 bv = *reinterpret_cast<const half8 *>(br + 8 * ks + 4 * (dl >> 5));
                    macc[s] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, macc[s], 0, 0, 0);
                }
                if constexpr (PG % 8 != 0) {
                    // the last 8 taps as a K = 16 step whose upper half is
                    // zero (lanes 32-63), which takes fewer passes than K = 8
                    const bool lo = dl < 32;
                    const u32x4 az = *reinterpret_cast<const u32x4 *>(ar + PG - 4);
                    const u32x4 bz = *reinterpret_cast<const u32x4 *>(br + PG - 4);
                    const half8 av = __builtin_bit_cast(half8, lo ? az : u32x4{0u, 0u, 0u, 0u});
                    const half8 bv = __builtin_bit_cast(half8, lo ? bz : u32x4{0u, 0u, 0u, 0u});
                    macc[s] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, macc[s], 0, 0, 0);
                }
            }
        }
        // the next window row's loads go out once this row's cost windows
        // and matrix operands are read (their LDS writes then meet the
        // column sums rather than those reads)
        if (yy < yt + r) issue(which ^ 1, yy + 1);
        // numerators: target row c_k = pxw + 2k - base + DT for k from
        // -(M-1); output (k, j) uses c_(k-j)
        uint32_t tv[AW_NX + M][NT];
        // one per-lane base; the rows of the columns at immediate offsets
        const uint32_t *tb = buf + T::T_OFF + opaque((pxw - 2 * (M - 1) - base + DT) * PG);
        auto load_t = [&](int kk) {
            const uint32_t *tr = tb + 2 * kk * PG;
#pragma unroll
            for (int g = 0; g < NT / 4; g++) {
                const uint4 v = reinterpret_cast<const uint4 *>(tr)[g];
                tv[kk][4 * g] = v.x;
                tv[kk][4 * g + 1] = v.y;
                tv[kk][4 * g + 2] = v.z;
                tv[kk][4 * g + 3] = v.w;
            }
#pragma unroll
            for (int i = 4 * (NT / 4); i < NT; i++) tv[kk][i] = tr[i];
        };
        // reference row of column k (its NT words)
        uint32_t lv[AW_NX][NT];
        auto load_l = [&](int k) {
            const uint4 *lw = reinterpret_cast<const uint4 *>(Lt + (pxw + 2 * k) * PG);
#pragma unroll
            for (int g = 0; g < NT / 4; g++) {
                const uint4 v = lw[g];
                lv[k][4 * g] = v.x;
                lv[k][4 * g + 1] = v.y;
                lv[k][4 * g + 2] = v.z;
                lv[k][4 * g + 3] = v.w;
            }
#pragma unroll
            for (int i = 4 * (NT / 4); i < NT; i++) lv[k][i] = Lt[(pxw + 2 * k) * PG + i];
        };
        // software pipeline: column k+1's target row (M > 1) and reference
        // row (M = 2) are read while column k is summed, as registers permit
        // (the scheduling barrier keeps the reads ahead of the sums)
        constexpr bool PRE_T = M > 1, PRE_L = M == 2;
#pragma unroll
        for (int kk = 0; kk < (PRE_T ? M : M - 1); kk++) load_t(kk);
        if (PRE_L) load_l(0);
#pragma unroll
        for (int k = 0; k < AW_NX; k++) {
            if (!PRE_T) load_t(k + M - 1);
            if (!PRE_L) load_l(k);
            if (PRE_T && k + 1 < AW_NX) {
                load_t(k + M);
                if (PRE_L) load_l(k + 1);
            }
            if (PRE_T) __builtin_amdgcn_sched_barrier(0);
            const uint32_t *l32 = lv[k];
#pragma unroll
            for (int j = 0; j < M; j++) {
                const uint32_t *t32 = tv[k + M - 1 - j];
                h2 p[NT];
#pragma unroll
                for (int i = 0; i < NT; i++) p[i] = __builtin_bit_cast(h2, l32[i]) * __builtin_bit_cast(h2, t32[i]);
                float sn = num[k][j];
#pragma unroll
                for (int i = 0; i < NT; i++)
                    sn = __builtin_amdgcn_fdot2(p[i], __builtin_bit_cast(h2, cw[j][k + i]), sn, false);
                num[k][j] = sn;
            }
        }
    };

    // One barrier per window row: each wave's reads of the previous row have
    // returned (lgkmcnt 0) and its loads of this row have landed (vmcnt 0);
    // past the barrier the row is complete in its buffer and the other
    // buffer is free for the next row's loads, which fly during this row's
    // sums.
    const int ylast = yt + r;
    auto step = [&](int cur, int yy) {
        __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
        wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        compute(cur, yy);
    };
    issue(0, yt - r);
    for (int yy = yt - r; yy <= ylast; yy++) step((yy - yt + r) & 1, yy);

    // The denominators through a 64 x QW float matrix (odd pitch), the
    // quotients into registers, then through LDS so that each disparity row
    // of 64 pixels is stored contiguously.
    float *dm = reinterpret_cast<float *>(lbuf);
    __syncthreads();
#pragma unroll
    for (int s = 0; s < T::NSLOT; s++) {
        const int b = wave + T::NWAVE * s;
        if (b < T::NBLK) {
            const int mt = b / (2 * M + 1), nt = mt + b % (2 * M + 1);
#pragma unroll
            for (int q = 0; q < 16; q++) {
                const int i = 32 * mt + 8 * (q / 4) + 4 * (dl >> 5) + (q % 4), c = 32 * nt + (dl & 31);
                dm[i * DMP + c] = macc[s][q];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < AW_NX; k++)
#pragma unroll
        for (int j = 0; j < M; j++) {
            const int px = pxw + 2 * k;
            num[k][j] = num[k][j] / dm[px * DMP + px - (base + 2 * j) + DT];
        }
    __syncthreads();
    float *tile = reinterpret_cast<float *>(lbuf);
#pragma unroll
    for (int k = 0; k < AW_NX; k++)
#pragma unroll
        for (int j = 0; j < M; j++) tile[(base + 2 * j) * 65 + pxw + 2 * k] = num[k][j];
    __syncthreads();
    for (int i = tid; i < DT * 64; i += T::NTHR) {
        const int dd = i >> 6, px = i & 63;
        const int d = d0 + dd, x = x0 + px;
        if (d < D && x < W) stream_store(&out[(size_t)d * pl + (size_t)yt * W + x], tile[dd * 65 + px]);
    }
}

// Disparities per lane for the FP16-volume kernel: the fewest lane-column
// passes, weighing an output of M = 2 at 0.6 and M = 3 at 0.45 of M = 1
// (their LDS bytes per output).
inline int aw_half_m(int D) {
    int best = 1;
    double bc = 1e30;
    for (int m = 1; m <= 3; m++) {
        const double c = (double)div_up(D, 64 * m) * 64 * m * (m == 1 ? 1.0 : m == 2 ? 0.6 : 0.45);
        if (c < bc) bc = c, best = m;
    }
    return best;
}

template <int R, int M>
int aw_half_band_launch(const void *wl, const void *wt, const uint16_t *volh, float *out, int W, int H, int D, int Wp,
                        int y0, int rows, int r, hipStream_t s) {
    using T = AwHalf<R, M>;
    const int nxb = div_up(W, T::TXP), nyb = rows, ndb = div_up(D, T::DT);
    k_aw_band_h<R, M><<<nxb * nyb * ndb, T::NTHR, 0, s>>>(static_cast<const uint32_t *>(wl),
                                                         static_cast<const uint32_t *>(wt), volh, out, W, H, D, Wp,
                                                         y0, r, nxb, nyb, ndb);
    RTS_LAUNCHED();
    return RTS_OK;
}

template <bool HALF, int R>
int aw_band_launch(const void *wl, const void *wt, const float *vol, float *out, int W, int H, int D, int Wp, int y0,
                   int rows, int r, hipStream_t s) {
    // FP16: the direct-load kernel, whose per-lane cost offsets span 64
    // planes in 32 bits
    if (HALF && 64.0 * (double)W * (double)H < 4294967296.0) {
        // one tile row per block: two blocks (two double buffers) share a CU
        // (the cost rows as 16-byte words where their columns x0-R+4i are
        // 16-byte aligned in every plane and across the wrap)
        constexpr int NY = 1;
        using TD = AwDma<R, NY, false>;
        const int nxb = div_up(W, TD::TXP), nyb = div_up(rows, NY), ndb = div_up(D, 64);
        const uint32_t *wlp = static_cast<const uint32_t *>(wl), *wtp = static_cast<const uint32_t *>(wt);
        if (R % 4 == 0 && W % 4 == 0 && (reinterpret_cast<uintptr_t>(vol) & 15) == 0)
            k_aw_band_dma<R, NY, true><<<nxb * nyb * ndb, TD::NTHR, 0, s>>>(wlp, wtp, vol, out, W, H, D, Wp, y0,
                                                                          rows, r, nxb, nyb, ndb);
        else
            k_aw_band_dma<R, NY, false><<<nxb * nyb * ndb, TD::NTHR, 0, s>>>(wlp, wtp, vol, out, W, H, D, Wp, y0,
                                                                           rows, r, nxb, nyb, ndb);
        RTS_LAUNCHED();
        return RTS_OK;
    }
    using T = AwTile<HALF, R>;
    constexpr size_t lds_bytes = sizeof(uint32_t) * T::WORDS;
    static bool attr_set = false;   // the attribute is per function; setting it twice is harmless
    if (lds_bytes > 65536 && !attr_set) {
        RTS_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_aw_band<HALF, R>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
        attr_set = true;
    }
    const int nxb = div_up(W, T::TXP), nyb = div_up(rows, AW_NY), ndb = div_up(D, 64);
    k_aw_band<HALF, R><<<nxb * nyb * ndb, T::NTHR, lds_bytes, s>>>(static_cast<const uint32_t *>(wl),
                                                                   static_cast<const uint32_t *>(wt), vol, out, W, H,
                                                                   D, Wp, y0, rows, r, nxb, nyb, ndb);
    RTS_LAUNCHED();
    return RTS_OK;
}

struct AwPlan {
    int R, NT, PG, Wp, band;
    size_t view_bytes;   // workspace of one view for one band
};

AwPlan aw_plan(int W, int H, int r, bool half) {
    AwPlan p;
    p.R = aw_rmax_for(r);
    p.NT = aw_ntap(p.R, half);
    p.PG = (p.NT + 3) & ~3;
    p.Wp = div_up(W, AW_TX) * AW_TX;
    const size_t row = (size_t)(2 * r + 1) * p.PG * p.Wp * 4;   // one view, one image row
    int band = (int)std::min<size_t>((size_t)H, std::max<size_t>(AW_NY, AW_WS_CAP / (2 * row)));
    band = std::min(band, 4 * 65535 / (2 * r + 1));   // grid.y of the weight kernel
    band -= band % AW_NY;
    if (band < AW_NY) band = AW_NY;
    // bands of equal height (no short last band with a small grid)
    if (band < H) {
        band = div_up(H, div_up(H, band));
        band += (AW_NY - band % AW_NY) % AW_NY;
    }
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
        const int nrow = rows * (2 * r + 1);
        const dim3 grid(pl.Wp / 64, div_up(nrow, 4));
        const float kc = (float)(-1.4426950408889634 / (double)gc);
        k_aw_weights_band<HALF, R>
            <<<grid, 256, 0, s>>>(labL, static_cast<uint32_t *>(wl), W, H, pl.Wp, y0, nrow, r, kc, gp);
        RTS_LAUNCHED();
        k_aw_weights_band<HALF, R>
            <<<grid, 256, 0, s>>>(labR, static_cast<uint32_t *>(wt), W, H, pl.Wp, y0, nrow, r, kc, gp);
        RTS_LAUNCHED();
        if (int st = aw_band_launch<HALF, R>(wl, wt, vol, out, W, H, D, pl.Wp, y0, rows, r, s)) return st;
    }
    return RTS_OK;
}

template <int R, int M>
int aw_half_run(const uint8_t *labL, const uint8_t *labR, const uint16_t *volh, float *out, int W, int H, int D,
                int r, float gc, float gp, const AwPlan &pl, uint8_t *ws, hipStream_t s) {
    void *wl = ws, *wt = ws + pl.view_bytes;
    for (int y0 = 0; y0 < H; y0 += pl.band) {
        const int rows = std::min(pl.band, H - y0);
        const int nrow = rows * (2 * r + 1);
        const dim3 grid(pl.Wp / 64, div_up(nrow, 4));
        const float kc = (float)(-1.4426950408889634 / (double)gc);
        k_aw_weights_band<true, R>
            <<<grid, 256, 0, s>>>(labL, static_cast<uint32_t *>(wl), W, H, pl.Wp, y0, nrow, r, kc, gp);
        RTS_LAUNCHED();
        k_aw_weights_band<true, R>
            <<<grid, 256, 0, s>>>(labR, static_cast<uint32_t *>(wt), W, H, pl.Wp, y0, nrow, r, kc, gp);
        RTS_LAUNCHED();
        if (int st = aw_half_band_launch<R, M>(wl, wt, volh, out, W, H, D, pl.Wp, y0, rows, r, s)) return st;
    }
    return RTS_OK;
}

template <int M>
int aw_half_dispatch(const uint8_t *labL, const uint8_t *labR, const uint16_t *volh, float *out, int W, int H, int D,
                     int r, float gc, float gp, const AwPlan &pl, uint8_t *ws, hipStream_t s) {
    switch (pl.R) {
        case 3: return aw_half_run<3, M>(labL, labR, volh, out, W, H, D, r, gc, gp, pl, ws, s);
        case 5: return aw_half_run<5, M>(labL, labR, volh, out, W, H, D, r, gc, gp, pl, ws, s);
        case 8: return aw_half_run<8, M>(labL, labR, volh, out, W, H, D, r, gc, gp, pl, ws, s);
        default: return aw_half_run<16, M>(labL, labR, volh, out, W, H, D, r, gc, gp, pl, ws, s);
    }
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

extern "C" int rts_aw_half_supported(int W, int H, int D) {
    if (W < 1 || H < 1 || D < 1 || W % 8 != 0) return 0;
    // the per-lane cost offsets span a tile's 64*M planes in 32 bits
    return 2.0 * 64 * 3 * (double)W * (double)H < 4294967296.0 ? 1 : 0;
}

extern "C" rts_status rts_aw_half_ws(const uint8_t *lab_ref, const uint8_t *lab_tar, const void *vol_half, float *out,
                                     int W, int H, int D, int r, float gamma_c, float gamma_p, void *workspace,
                                     size_t workspace_bytes, void *stream) {
    if (int e = aw_check(static_cast<const float *>(vol_half), out, W, H, D, r, RTS_AW_FP16)) return e;
    RTS_CHECK(rts_aw_half_supported(W, H, D), "FP16-volume AW needs a width that is a multiple of 8 and "
              "384*W*H below 2^32 (W %d, H %d)", W, H);
    RTS_CHECK((reinterpret_cast<uintptr_t>(vol_half) & 15) == 0, "the FP16 volume must be 16-byte aligned");
    const AwPlan pl = aw_plan(W, H, r, true);
    RTS_CHECK(workspace && workspace_bytes >= 2 * pl.view_bytes, "AW workspace of %zu bytes is below the %zu needed",
              workspace_bytes, 2 * pl.view_bytes);
    hipStream_t s = S(stream);
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    const uint16_t *vh = static_cast<const uint16_t *>(vol_half);
    switch (aw_half_m(D)) {
        case 1: return aw_half_dispatch<1>(lab_ref, lab_tar, vh, out, W, H, D, r, gamma_c, gamma_p, pl, ws, s);
        case 2: return aw_half_dispatch<2>(lab_ref, lab_tar, vh, out, W, H, D, r, gamma_c, gamma_p, pl, ws, s);
        default: return aw_half_dispatch<3>(lab_ref, lab_tar, vh, out, W, H, D, r, gamma_c, gamma_p, pl, ws, s);
    }
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
