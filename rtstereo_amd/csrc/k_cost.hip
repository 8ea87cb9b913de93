// Matching-cost kernels: AD (gray, float, colour), census + Hamming, rank,
// LoG convolution, NCC and the exponential cost combination.
//
// Every volume kernel writes the disparity-major layout [D][H][W]: a block
// covers 256 consecutive x of one row and walks a chunk of disparities, so
// each store instruction writes 256 contiguous floats.
#include "rts_internal.h"

#pragma clang fp contract(off)

namespace rts {

constexpr int XB = 256;   // threads per block along x
constexpr int DCH = 16;   // disparities per block

enum { AD_GRAY = 0, AD_FLOAT = 1, AD_COLOR = 2 };

// seq/PixelBasedCost.cpp:13-158.  dir 0: tar at x-d (L2R); 1: tar at x+d (R2L).
// OUT float, or _Float16 (each cost rounded to nearest) for FP16 consumers.
template <int KIND, typename OUT = float>
__global__ void __launch_bounds__(XB) k_ad(const void *__restrict__ ref_, const void *__restrict__ tar_,
                                           OUT *__restrict__ vol, int W, int H, int D, int dir, int use_trunc,
                                           float trunc, int norm) {
    const int x = blockIdx.x * XB + threadIdx.x;
    const int y = blockIdx.y;
    const int d0 = blockIdx.z * DCH;
    if (x >= W) return;
    const size_t row = (size_t)y * W;
    const size_t pl = plane(W, H);
    const int dend = min(D, d0 + DCH);
    if (KIND == AD_GRAY) {
        const uint8_t *ref = (const uint8_t *)ref_, *tar = (const uint8_t *)tar_;
        const int r = ref[row + x];
        for (int d = d0; d < dend; d++) {
            const int xt = wrap1(dir ? x + d : x - d, W);
            const int t = r - (int)tar[row + xt];
            vol[d * pl + row + x] = (OUT)(float)(t < 0 ? -t : t);
        }
    } else if (KIND == AD_FLOAT) {
        const float *ref = (const float *)ref_, *tar = (const float *)tar_;
        const float r = ref[row + x];
        for (int d = d0; d < dend; d++) {
            const int xt = wrap1(dir ? x + d : x - d, W);
            float v = r - tar[row + xt];
            if (v < 0) v = -v;
            vol[d * pl + row + x] = (OUT)v;
        }
    } else {
        const uint8_t *ref = (const uint8_t *)ref_, *tar = (const uint8_t *)tar_;
        const uint32_t r = pack_rgb(ref + 3 * (row + x));
        const float third = (float)1 / (float)3;
        for (int d = d0; d < dend; d++) {
            const int xt = wrap1(dir ? x + d : x - d, W);
            const int s = (int)sad3(r, pack_rgb(tar + 3 * (row + xt)));
            float v = use_trunc ? (((float)s > trunc) ? trunc : (float)s) : (float)s;
            if (norm) v = third * v;
            vol[d * pl + row + x] = (OUT)v;
        }
    }
}

// The colour AD of k_ad<AD_COLOR>, eight pixels per thread: a block stages
// one row of both views (packed colour) in LDS and each thread stores its
// eight costs of a disparity as 16-byte words (W a multiple of 8): one word
// of halves (OUT _Float16, each float cost rounded to nearest) or two of
// floats.  The same float arithmetic as k_ad.
constexpr int ADH_DCH = 32;   // disparities per block
template <typename OUT>
__global__ void __launch_bounds__(256) k_ad_color_h8(const uint8_t *__restrict__ ref, const uint8_t *__restrict__ tar,
                                                     OUT *__restrict__ vol, int W, int H, int D, int dir,
                                                     int use_trunc, float trunc, int norm) {
    // ref and tar rows, one padding word after every 8 (lanes reading at a
    // stride of 8 pixels then meet no bank conflicts)
    extern __shared__ uint32_t rows[];
    const int y = blockIdx.x, d0 = blockIdx.y * ADH_DCH;
    const int PW = W + W / 8;
    const size_t row = (size_t)y * W;
    for (int i = threadIdx.x; i < W; i += 256) {
        rows[i + (i >> 3)] = pack_rgb(ref + 3 * (row + i));
        rows[PW + i + (i >> 3)] = pack_rgb(tar + 3 * (row + i));
    }
    __syncthreads();
    const float third = (float)1 / (float)3;
    const size_t pl = plane(W, H);
    const int dend = min(D, d0 + ADH_DCH);
    for (int x = 8 * threadIdx.x; x < W; x += 8 * 256) {
        uint32_t r[8];
#pragma unroll
        for (int i = 0; i < 8; i++) r[i] = rows[x + (x >> 3) + i];
        for (int d = d0; d < dend; d++) {
            float c[8];
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int xt = wrap1(dir ? x + i + d : x + i - d, W);
                const int sd = (int)sad3(r[i], rows[PW + xt + (xt >> 3)]);
                float v = use_trunc ? (((float)sd > trunc) ? trunc : (float)sd) : (float)sd;
                if (norm) v = third * v;
                c[i] = v;
            }
            typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
            u32x4_t *dst = reinterpret_cast<u32x4_t *>(vol + (size_t)d * pl + row + x);
            if constexpr (sizeof(OUT) == 2) {
                uint32_t o[4];
#pragma unroll
                for (int i = 0; i < 8; i += 2)
                    o[i / 2] = (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)c[i]) |
                               (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)c[i + 1]) << 16;
                __builtin_nontemporal_store(u32x4_t{o[0], o[1], o[2], o[3]}, dst);
            } else {
                __builtin_nontemporal_store(u32x4_t{__float_as_uint(c[0]), __float_as_uint(c[1]), __float_as_uint(c[2]),
                                                    __float_as_uint(c[3])},
                                            dst);
                __builtin_nontemporal_store(u32x4_t{__float_as_uint(c[4]), __float_as_uint(c[5]), __float_as_uint(c[6]),
                                                    __float_as_uint(c[7])},
                                            dst + 1);
            }
        }
    }
}

// seq/CensusTransform.cpp:11-37.  Entry k = (dy+hr)*(2wr+1) + (dx+wr) is set
// when the centre is strictly brighter than the wrapped neighbour; entries
// are packed little-endian into 64-bit words, word-major [nw][H][W].
__global__ void __launch_bounds__(256) k_census(const uint8_t *__restrict__ img, uint64_t *__restrict__ words, int W,
                                                int H, int hr, int wr) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const uint8_t c = img[(size_t)y * W + x];
    const size_t pl = plane(W, H);
    uint64_t acc = 0;
    int k = 0, w = 0;
    for (int dy = -hr; dy <= hr; dy++) {
        const size_t rr = (size_t)wrap1(y + dy, H) * W;
        for (int dx = -wr; dx <= wr; dx++) {
            acc |= (uint64_t)(c > img[rr + wrap1(x + dx, W)]) << k;
            if (++k == 64) {
                words[w * pl + (size_t)y * W + x] = acc;
                acc = 0;
                k = 0;
                w++;
            }
        }
    }
    if (k) words[w * pl + (size_t)y * W + x] = acc;
}

// The same with the block's image rows staged in LDS (radii up to 8).
constexpr int CT_R = 8;
__global__ void __launch_bounds__(256) k_census_tile(const uint8_t *__restrict__ img, uint64_t *__restrict__ words,
                                                     int W, int H, int hr, int wr) {
    __shared__ uint8_t tile[4 + 2 * CT_R][64 + 2 * CT_R];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    const int th = 4 + 2 * hr, tw = 64 + 2 * wr;
    for (int i = threadIdx.y * 64 + threadIdx.x; i < th * tw; i += 256) {
        const int r = i / tw, c = i - r * tw;
        tile[r][c] = img[(size_t)wrapn(y0 - hr + r, H) * W + wrapn(x0 - wr + c, W)];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= W || y >= H) return;
    const int tx = threadIdx.x + wr, ty = threadIdx.y + hr;
    const uint8_t c = tile[ty][tx];
    const size_t pl = plane(W, H);
    uint64_t acc = 0;
    int k = 0, w = 0;
    for (int dy = -hr; dy <= hr; dy++)
        for (int dx = -wr; dx <= wr; dx++) {
            acc |= (uint64_t)(c > tile[ty + dy][tx + dx]) << k;
            if (++k == 64) {
                words[w * pl + (size_t)y * W + x] = acc;
                acc = 0;
                k = 0;
                w++;
            }
        }
    if (k) words[w * pl + (size_t)y * W + x] = acc;
}

// Census windows with wr <= 4 (the reference's 7x9 among them), four
// pixels per thread: the pixels x = 4i..4i+3 of a row see columns
// 4i-wr .. 4i+3+wr, all inside the three aligned dwords 4i-4 .. 4i+7 of the
// staged row, so every byte a comparison reads is a fixed byte of a register
// (byte-select compares).  Each comparison shifts its bit into the word with
// an add-with-carry, taken from the last window entry down so that entry k
// lands on bit k.  A block covers 256 columns x CW_TY rows (8 rows: twice
// the blocks of 16, 22 instead of 30 us a 1080p view); the staged rows
// wrap around the image edges like the loop form.
constexpr int CW_TY = 8;

template <int HR, int WR>
__global__ void __launch_bounds__(256) k_census_w4(const uint8_t *__restrict__ img, uint64_t *__restrict__ words,
                                                   int W, int H) {
    static_assert(WR <= 4 && (2 * HR + 1) * (2 * WR + 1) <= 64, "window");
    constexpr int TR = CW_TY + 2 * HR, TWD = 66;   // staged rows, dwords per row (columns x0-4 .. x0+259)
    __shared__ uint32_t tile[TR][TWD];
    const int x0 = blockIdx.x * 256, y0 = blockIdx.y * CW_TY;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const bool inner = (W % 4) == 0 && x0 >= 4 && x0 + 260 <= W;
    for (int i = tid; i < TR * TWD; i += 256) {
        const int r = i / TWD, c = i - r * TWD;
        const size_t ro = (size_t)wrapn(y0 - HR + r, H) * W;
        uint32_t v;
        if (inner) {
            v = *reinterpret_cast<const uint32_t *>(img + ro + x0 - 4 + 4 * c);
        } else {
            v = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) v |= (uint32_t)img[ro + wrapn(x0 - 4 + 4 * c + k, W)] << (8 * k);
        }
        tile[r][c] = v;
    }
    __syncthreads();
    const int tx = threadIdx.x, xb = x0 + 4 * tx;
    if (xb >= W) return;
    const size_t pl = plane(W, H);
    for (int ty = threadIdx.y; ty < CW_TY; ty += 4) {
        const int y = y0 + ty;
        if (y >= H) break;
        // dwords tx .. tx+2 of rows ty .. ty+2*HR: byte (4 + j + dx) of the
        // three is the neighbour dx of pixel j
        uint32_t rows[2 * HR + 1][3];
#pragma unroll
        for (int r = 0; r <= 2 * HR; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) rows[r][q] = tile[ty + r][tx + q];
        uint64_t out[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint32_t c = (rows[HR][(4 + j) >> 2] >> (8 * ((4 + j) & 3))) & 255u;
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int k = (2 * HR + 1) * (2 * WR + 1) - 1; k >= 0; k--) {
                const int dy = k / (2 * WR + 1), dx = k % (2 * WR + 1) - WR;
                const int bpos = 4 + j + dx;
                const uint32_t nb = (rows[dy][bpos >> 2] >> (8 * (bpos & 3))) & 255u;
                const uint32_t bit = c > nb;
                if (k >= 32)
                    hi = hi + hi + bit;
                else
                    lo = lo + lo + bit;
            }
            out[j] = (uint64_t)hi << 32 | lo;
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            if (xb + j < W) words[(size_t)y * W + xb + j] = out[j];
    }
}

__global__ void k_census_unpack(const uint64_t *__restrict__ words, uint8_t *__restrict__ bools, int P, int N) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    for (int k = 0; k < N; k++) bools[(size_t)N * p + k] = (uint8_t)((words[(size_t)(k >> 6) * P + p] >> (k & 63)) & 1u);
}

__global__ void k_census_pack(const uint8_t *__restrict__ bools, uint64_t *__restrict__ words, int P, int N) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int nw = (N + 63) / 64;
    for (int w = 0; w < nw; w++) {
        uint64_t acc = 0;
        for (int k = 64 * w; k < min(N, 64 * w + 64); k++)
            acc |= (uint64_t)(bools[(size_t)N * p + k] != 0) << (k & 63);
        words[(size_t)w * P + p] = acc;
    }
}

// seq/CensusTransform.cpp:39-62: popcount of the XOR of the descriptors.
__global__ void __launch_bounds__(XB) k_hamming(const uint64_t *__restrict__ cref, const uint64_t *__restrict__ ctar,
                                                float *__restrict__ vol, int W, int H, int D, int nw) {
    const int x = blockIdx.x * XB + threadIdx.x;
    const int y = blockIdx.y;
    const int d0 = blockIdx.z * DCH;
    if (x >= W) return;
    const size_t row = (size_t)y * W, pl = plane(W, H);
    const int dend = min(D, d0 + DCH);
    if (nw == 1) {
        const uint64_t r = cref[row + x];
        for (int d = d0; d < dend; d++)
            vol[d * pl + row + x] = (float)__popcll(r ^ ctar[row + wrap1(x - d, W)]);
    } else {
        for (int d = d0; d < dend; d++) {
            const int xt = wrap1(x - d, W);
            int n = 0;
            for (int w = 0; w < nw; w++) n += __popcll(cref[w * pl + row + x] ^ ctar[w * pl + row + xt]);
            vol[d * pl + row + x] = (float)n;
        }
    }
}

// seq/RankFilter.cpp:11-34: 1 + number of window pixels darker than the centre.
__global__ void __launch_bounds__(256) k_rank(const uint8_t *__restrict__ img, uint8_t *__restrict__ out, int W,
                                              int H, int m) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const uint8_t c = img[(size_t)y * W + x];
    int n = 1;
    for (int dy = -m; dy <= m; dy++) {
        const size_t rr = (size_t)wrap1(y + dy, H) * W;
        for (int dx = -m; dx <= m; dx++) n += img[rr + wrap1(x + dx, W)] < c;
    }
    out[(size_t)y * W + x] = (uint8_t)n;
}

struct ConvMask {
    float m[225];
};

// seq/ConvolveImage.cpp:11-34: float products added in row-major window
// order (contraction off keeps each product rounded), |.|, clamp to 255.
__global__ void __launch_bounds__(256) k_convolve(const uint8_t *__restrict__ img, float *__restrict__ out, int W,
                                                  int H, int hr, int wr, ConvMask mask) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= W || y >= H) return;
    const int mw = 2 * wr + 1;
    float acc = 0.0f;
    for (int dy = -hr; dy <= hr; dy++) {
        const size_t rr = (size_t)wrap1(y + dy, H) * W;
        for (int dx = -wr; dx <= wr; dx++) {
            const float prod = (float)img[rr + wrap1(x + dx, W)] * mask.m[(dy + hr) * mw + (dx + wr)];
            acc += prod;
        }
    }
    acc = acc > 0 ? acc : -acc;
    out[(size_t)y * W + x] = acc > 255 ? 255.0f : acc;
}

// seq/NCCDisparity.cpp:11-73 in the reference's own float order: window
// means as float sums / area, centred values, nom = sum a'b',
// den = sqrt(sum a'a'b'b') accumulated as ((a'a')b')b'.
__global__ void __launch_bounds__(256) k_ncc_seq(const uint8_t *__restrict__ ref, const uint8_t *__restrict__ tar,
                                                 float *__restrict__ vol, int W, int H, int D, int hr, int wr,
                                                 int keep_flat) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int d0 = blockIdx.z * DCH;
    if (x >= W || y >= H) return;
    const size_t pl = plane(W, H);
    const float area = (float)((2 * hr + 1) * (2 * wr + 1));
    float mr = 0.0f;
    for (int dy = -hr; dy <= hr; dy++) {
        const size_t rr = (size_t)wrap1(y + dy, H) * W;
        for (int dx = -wr; dx <= wr; dx++) mr += (float)ref[rr + wrap1(x + dx, W)];
    }
    mr = __fdiv_rn(mr, area);
    const int dend = min(D, d0 + DCH);
    for (int d = d0; d < dend; d++) {
        float mt = 0.0f;
        for (int dy = -hr; dy <= hr; dy++) {
            const size_t rr = (size_t)wrap1(y + dy, H) * W;
            for (int dx = -wr; dx <= wr; dx++) mt += (float)tar[rr + wrap1(x + dx - d, W)];
        }
        mt = __fdiv_rn(mt, area);
        float nom = 0.0f, den = 0.0f;
        for (int dy = -hr; dy <= hr; dy++) {
            const size_t rr = (size_t)wrap1(y + dy, H) * W;
            for (int dx = -wr; dx <= wr; dx++) {
                const float a = (float)ref[rr + wrap1(x + dx, W)] - mr;
                const float b = (float)tar[rr + wrap1(x + dx - d, W)] - mt;
                nom += a * b;
                den += ((a * a) * b) * b;
            }
        }
        den = sqrt_rn(den);
        const size_t o = d * pl + (size_t)y * W + x;
        if (den != 0.0f)
            vol[o] = __fdiv_rn(nom, den);
        else if (!keep_flat)
            vol[o] = 0.0f;
    }
}

// NCC from exact integer window sums.  With N the window area and a, b the
// ref and (shifted) tar values:
//   N*nom           = N*Sab - Sa*Sb
//   N^4*den^2       = N^4*Saabb - 2N^3(Sb*Saab + Sa*Sabb)
//                     + N^2(Sb^2*Saa + Sa^2*Sbb + 4*Sa*Sb*Sab) - 3N*Sa^2*Sb^2
// both exact (the second in 128-bit integers), so den == 0 is decided
// exactly; the quotient N*(N*nom)/sqrt(N^4 den^2) is formed in FP64.
__global__ void __launch_bounds__(256) k_ncc_int(const uint8_t *__restrict__ ref, const uint8_t *__restrict__ tar,
                                                 float *__restrict__ vol, int W, int H, int D, int hr, int wr,
                                                 int keep_flat) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    const int d0 = blockIdx.z * DCH;
    if (x >= W || y >= H) return;
    const size_t pl = plane(W, H);
    const long N = (long)(2 * hr + 1) * (2 * wr + 1);
    long Sa = 0, Saa = 0;
    for (int dy = -hr; dy <= hr; dy++) {
        const size_t rr = (size_t)wrap1(y + dy, H) * W;
        for (int dx = -wr; dx <= wr; dx++) {
            const long a = ref[rr + wrap1(x + dx, W)];
            Sa += a;
            Saa += a * a;
        }
    }
    const int dend = min(D, d0 + DCH);
    for (int d = d0; d < dend; d++) {
        long Sb = 0, Sbb = 0, Sab = 0, Saab = 0, Sabb = 0;
        unsigned long Saabb = 0;
        for (int dy = -hr; dy <= hr; dy++) {
            const size_t rr = (size_t)wrap1(y + dy, H) * W;
            for (int dx = -wr; dx <= wr; dx++) {
                const long a = ref[rr + wrap1(x + dx, W)];
                const long b = tar[rr + wrap1(x + dx - d, W)];
                const long ab = a * b;
                Sb += b;
                Sbb += b * b;
                Sab += ab;
                Saab += ab * a;
                Sabb += ab * b;
                Saabb += (unsigned long)(ab * ab);
            }
        }
        const __int128 n = N;
        const __int128 X = n * n * n * n * (__int128)Saabb - 2 * n * n * n * ((__int128)Sb * Saab + (__int128)Sa * Sabb) +
                           n * n * ((__int128)Sb * Sb * Saa + (__int128)Sa * Sa * Sbb + 4 * (__int128)Sa * Sb * Sab) -
                           3 * n * (__int128)Sa * Sa * Sb * Sb;
        const long nomN = N * Sab - Sa * Sb;
        const size_t o = d * pl + (size_t)y * W + x;
        if (X > 0)
            vol[o] = (float)((double)N * (double)nomN / sqrt((double)X));
        else if (!keep_flat)
            vol[o] = 0.0f;
    }
}

// Three-term cost of the batched workload, in one pass: per (x, y, d)
//   c = (1-exp(-|rankL(x)-rankR(x-d)|/l1)) + (1-exp(-|logL(x)-logR(x-d)|/l2))
//     + (1-exp(-(1-ncc)/l3)),
// ncc from the same exact integer window sums as k_ncc_int (0 where the
// window is flat).  Block: 64 x 4 pixels x NC_DCH disparities; the gray
// rows the block's windows touch are staged in LDS.
constexpr int NC_RMAX = 5;
constexpr int NC_TH = 4 + 2 * NC_RMAX;   // staged rows

// The k_ncc_int quotient with the denominator factored as N^4 den^2 = N * Z,
//   Z = N*(N^2 Saabb - 2N(Sb Saab + Sa Sabb) + Sb^2 Saa + Sa^2 Sbb + 4 Sa Sb Sab) - 3 Sa^2 Sb^2,
// Z = sum over taps of (N a - Sa)^2 (N b - Sb)^2 / N <= (255 N)^4: for
// windows of up to 121 taps every partial term stays below 2^63, so Z is
// exact in 64-bit integers and the flat test Z == 0 is exact.
__device__ __forceinline__ float ncc_exact(long N, long Sa, long Saa, long Sb, long Sbb, long Sab, long Saab, long Sabb,
                                           unsigned long Saabb) {
    const long Y = N * N * (long)Saabb - 2 * N * (Sb * Saab + Sa * Sabb) + Sb * Sb * Saa + Sa * Sa * Sbb +
                   4 * Sa * Sb * Sab;
    const long Z = N * Y - 3 * (Sa * Sb) * (Sa * Sb);
    const long nomN = N * Sab - Sa * Sb;
    return Z > 0 ? (float)((double)N * (double)nomN / sqrt((double)N * (double)Z)) : 0.0f;
}

template <int NC_DCH>
__global__ void __launch_bounds__(256) k_cost_rank_log_ncc(const uint8_t *__restrict__ gl, const uint8_t *__restrict__ gr,
                                                           const uint8_t *__restrict__ rkl,
                                                           const uint8_t *__restrict__ rkr,
                                                           const float *__restrict__ lgl, const float *__restrict__ lgr,
                                                           float *__restrict__ vol, int W, int H, int D, int hr,
                                                           int wr, float l1, float l2, float l3) {
    constexpr int NC_TW = 64 + 2 * NC_RMAX + NC_DCH;   // staged columns (tar)
    __shared__ uint8_t ta[NC_TH][64 + 2 * NC_RMAX], tb[NC_TH][NC_TW];
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4, d0 = blockIdx.z * NC_DCH;
    const int tid = threadIdx.y * 64 + threadIdx.x;
    const int rows = 4 + 2 * hr, acols = 64 + 2 * wr, bcols = 64 + 2 * wr + NC_DCH - 1;
    for (int i = tid; i < rows * acols; i += 256) {
        const int r = i / acols, c = i - r * acols;
        ta[r][c] = gl[(size_t)wrapn(y0 - hr + r, H) * W + wrapn(x0 - wr + c, W)];
    }
    for (int i = tid; i < rows * bcols; i += 256) {
        const int r = i / bcols, c = i - r * bcols;
        tb[r][c] = gr[(size_t)wrapn(y0 - hr + r, H) * W + wrapn(x0 - wr - d0 - (NC_DCH - 1) + c, W)];
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= W || y >= H) return;
    const size_t pl = plane(W, H);
    const long N = (long)(2 * hr + 1) * (2 * wr + 1);
    const int ty = threadIdx.y, tx = threadIdx.x;
    int Sa = 0, Saa = 0;
    for (int dy = 0; dy <= 2 * hr; dy++)
        for (int dx = 0; dx <= 2 * wr; dx++) {
            const int a = ta[ty + dy][tx + dx];
            Sa += a;
            Saa += a * a;
        }
    const float rl = rkl[(size_t)y * W + x], ll = lgl[(size_t)y * W + x];
    const int dend = min(D, d0 + NC_DCH);
    for (int d = d0; d < dend; d++) {
        // tar column of window tap dx for disparity d: x + dx - wr - d
        const int bc = tx + (NC_DCH - 1) - (d - d0);
        int Sb = 0, Sbb = 0, Sab = 0, Saab = 0, Sabb = 0;
        unsigned long Saabb = 0;
        for (int dy = 0; dy <= 2 * hr; dy++)
            for (int dx = 0; dx <= 2 * wr; dx++) {
                const int a = ta[ty + dy][tx + dx], b = tb[ty + dy][bc + dx];
                const int ab = a * b;
                Sb += b;
                Sbb += b * b;
                Sab += ab;
                Saab += ab * a;
                Sabb += ab * b;
                Saabb += (unsigned long)((unsigned)ab * (unsigned)ab);
            }
        const float s = ncc_exact(N, Sa, Saa, Sb, Sbb, Sab, Saab, Sabb, Saabb);
        const int xt = wrap1(x - d, W);
        const float cr = fabsf(rl - (float)rkr[(size_t)y * W + xt]);
        const float cl = fabsf(ll - lgr[(size_t)y * W + xt]);
        const float t1 = 1.0f - expf(__fdiv_rn(-cr, l1));
        const float t2 = 1.0f - expf(__fdiv_rn(-cl, l2));
        const float t3 = 1.0f - expf(__fdiv_rn(-(1.0f - s), l3));
        vol[d * pl + (size_t)y * W + x] = (t1 + t2) + t3;
    }
}

// The same cost for a (2HR+1) x (2WR+1) NCC window by a row sweep.  Block:
// 64 columns x a segment of NS_SEG output rows, four waves each on its own
// disparity, DCH/4 disparity rounds.  Per row each wave stages the columns
// its lanes' windows touch in LDS (with the view records: each column's
// products ab, a^2b, ab^2, a^2b^2, formed once; without: the (a, b) pair)
// and forms the row's horizontal window sums of a, a^2, b, b^2, ab, a^2b,
// ab^2, a^2b^2 (with the records only the last four);
// the last 2HR+1 of them stay in a register ring, so the vertical sums move
// by one add and one subtract per row.  The window sums are the exact
// integers of k_cost_rank_log_ncc; the quotient differs from that kernel's
// in its last bits (double intermediates and a hardware reciprocal square
// root here, an exact integer Z and a double square root there).
constexpr int NS_SEG = 64;

struct NccSums {
    int sa, saa, sb, sbb, sab, saab, sabb;
    unsigned long saabb;
};

__device__ __forceinline__ void ncc_add(NccSums &s, const NccSums &h, const NccSums &o) {
    s.sa += h.sa - o.sa;
    s.saa += h.saa - o.saa;
    s.sb += h.sb - o.sb;
    s.sbb += h.sbb - o.sbb;
    s.sab += h.sab - o.sab;
    s.saab += h.saab - o.saab;
    s.sabb += h.sabb - o.sabb;
    s.saabb += h.saabb - o.saabb;
}

// The NCC value from the window sums in double arithmetic (for a 7x7 window
// every product in Y and nomN is below 2^53 and exact; N*Y and (Sa*Sb)^2 may
// round, ~1e-16 relative) and a float quotient with the hardware reciprocal
// square root (~1e-7 relative).  A window with a flat side has nomN == 0
// exactly, so it gives 0 either way.
__device__ __forceinline__ float ncc_fast(long N, long Sa, long Saa, long Sb, long Sbb, long Sab, long Saab, long Sabb,
                                          unsigned long Saabb) {
    const double n = (double)N, a = (double)Sa, aa = (double)Saa, b = (double)Sb, bb = (double)Sbb;
    const double ab = (double)Sab, aab = (double)Saab, abb = (double)Sabb, aabb = (double)Saabb;
    const double Y = n * n * aabb - 2.0 * n * (b * aab + a * abb) + b * b * aa + a * a * bb + 4.0 * a * b * ab;
    const double Z = n * Y - 3.0 * (a * b) * (a * b);
    const double nomN = n * ab - a * b;
    return Z > 0 ? (float)n * (float)nomN * __builtin_amdgcn_rsqf((float)n * (float)Z) : 0.0f;
}

// Per-pixel record of one view for the row-sweep cost: the window sums of
// the gray level and of its square over the (2HR+1) x (2WR+1) window
// (wrapped borders, exact integers), the rank and the LoG value's bits.
// They do not depend on the disparity, so the sweep reads them instead of
// forming them for every disparity.
template <int HR, int WR>
__global__ void __launch_bounds__(256) k_ncc_records(const uint8_t *__restrict__ g, const uint8_t *__restrict__ rk,
                                                     const float *__restrict__ lg, uint4 *__restrict__ rec, int W,
                                                     int H) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= W || y >= H) return;
    const bool inner = x - WR >= 0 && x + WR < W;
    unsigned s = 0, ss = 0;
    for (int dy = -HR; dy <= HR; dy++) {
        const uint8_t *row = g + (size_t)wrapn(y + dy, H) * W;
#pragma unroll
        for (int dx = -WR; dx <= WR; dx++) {
            const unsigned a = row[inner ? x + dx : wrapn(x + dx, W)];
            s += a;
            ss += a * a;
        }
    }
    const size_t o = (size_t)y * W + x;
    rec[o] = make_uint4(s, ss, rk[o], __float_as_uint(lg[o]));
}

// PRE: the gray window sums, ranks and LoG values come from the two views'
// records (k_ncc_records), so each tap forms only the four sums with
// products of both views; otherwise all eight sums are formed here.  The
// integers reaching the quotient, and so the outputs, are the same.
template <int HR, int WR, int DCH, bool PRE>
__global__ void __launch_bounds__(256) k_cost_rln_sweep(const uint8_t *__restrict__ gl, const uint8_t *__restrict__ gr,
                                                        const uint8_t *__restrict__ rkl, const uint8_t *__restrict__ rkr,
                                                        const float *__restrict__ lgl, const float *__restrict__ lgr,
                                                        const uint4 *__restrict__ recl, const uint4 *__restrict__ recr,
                                                        float *__restrict__ vol, int W, int H, int D, float l1,
                                                        float l2, float l3) {
    constexpr int E = 2 * HR + 1, CW = 64 + 2 * WR;
    constexpr long N = (long)E * (2 * WR + 1);
    // PRE: a column's four products with both views (ab, a^2 b, a b^2,
    // a^2 b^2: all below 2^32) are formed once, where the column is
    // staged, and the taps read them; otherwise the (a, b) pair is staged
    __shared__ uint16_t buf[PRE ? 1 : 4][PRE ? 1 : CW];
    __shared__ uint4 pbuf[PRE ? 4 : 1][PRE ? CW : 1];
    const float k1 = -1.4426950408889634f / l1, k2 = -1.4426950408889634f / l2, k3 = -1.4426950408889634f / l3;
    const int x0 = blockIdx.x * 64, s0 = blockIdx.y * NS_SEG, s1 = min(H, s0 + NS_SEG);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = x0 + lane;
    const size_t pl = plane(W, H);
    const int nr = (s1 - s0) + 2 * HR;   // rows swept (the same for every wave)
    for (int round = 0; round < DCH / 4; round++) {
        const int d = blockIdx.z * DCH + round * 4 + w;
        const bool dok = d < D;
        const int dd = dok ? d : 0;
        NccSums ring[E], S = {};
#pragma unroll
        for (int k = 0; k < E; k++) ring[k] = NccSums{};
        for (int base = 0; base < nr; base += E) {
#pragma unroll
            for (int k = 0; k < E; k++) {
                const int i = base + k;
                if (i >= nr) break;
                const int y = wrap1(s0 - HR + i, H);   // 2HR+1 <= H: at most one period out
                // Each wave works on its own LDS row: a wave-level fence and
                // barrier order its previous reads before these writes, and
                // these writes before the reads below.
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                for (int c = lane; c < CW; c += 64) {
                    int xa = x0 - WR + c, xb = xa - dd;
                    xa = xa < 0 ? xa + W : (xa >= W ? xa - W : xa);
                    xb = xb < 0 ? xb + W : (xb >= W ? xb - W : xb);
                    if (xa < 0 || xa >= W) xa = wrapn(xa, W);   // images narrower than the window reach
                    if (xb < 0 || xb >= W) xb = wrapn(xb, W);
                    const unsigned a = gl[(size_t)y * W + xa], b = gr[(size_t)y * W + xb];
                    if constexpr (PRE) {
                        const unsigned ab = a * b;
                        pbuf[w][c] = make_uint4(ab, a * ab, b * ab, ab * ab);
                    } else {
                        buf[w][c] = (uint16_t)(a | (b << 8));
                    }
                }
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                NccSums h = {};
                if constexpr (PRE) {
#pragma unroll
                    for (int dx = 0; dx <= 2 * WR; dx++) {
                        const uint4 q = pbuf[w][lane + dx];
                        h.sab += (int)q.x;
                        h.saab += (int)q.y;
                        h.sabb += (int)q.z;
                        h.saabb += (unsigned long)q.w;
                    }
                }
#pragma unroll
                for (int dx = 0; dx <= 2 * WR && !PRE; dx++) {
                    const int v = buf[w][lane + dx];
                    const int a = v & 255, b = v >> 8, ab = a * b;
                    if constexpr (!PRE) {
                        h.sa += a;
                        h.saa += a * a;
                        h.sb += b;
                        h.sbb += b * b;
                    }
                    h.sab += ab;
                    h.saab += a * ab;
                    h.sabb += b * ab;
                    h.saabb += (unsigned long)((unsigned)ab * (unsigned)ab);
                }
                ncc_add(S, h, ring[k]);
                ring[k] = h;
                if (i >= 2 * HR && dok && x < W) {
                    const int yo = s0 + i - 2 * HR;
                    const size_t o = (size_t)yo * W;
                    const int xt = wrap1(x - d, W);
                    float sn, cr, cl;
                    if constexpr (PRE) {
                        const uint4 ra = recl[o + x], rb = recr[o + xt];
                        sn = ncc_fast(N, ra.x, ra.y, rb.x, rb.y, S.sab, S.saab, S.sabb, S.saabb);
                        cr = fabsf((float)ra.z - (float)rb.z);
                        cl = fabsf(__uint_as_float(ra.w) - __uint_as_float(rb.w));
                    } else {
                        sn = ncc_fast(N, S.sa, S.saa, S.sb, S.sbb, S.sab, S.saab, S.sabb, S.saabb);
                        cr = fabsf((float)rkl[o + x] - (float)rkr[o + xt]);
                        cl = fabsf(lgl[o + x] - lgr[o + xt]);
                    }
                    // exp(-c/l) = exp2(c * k), k = -log2(e)/l (hardware exp2)
                    const float t1 = 1.0f - __builtin_amdgcn_exp2f(cr * k1);
                    const float t2 = 1.0f - __builtin_amdgcn_exp2f(cl * k2);
                    const float t3 = 1.0f - __builtin_amdgcn_exp2f((1.0f - sn) * k3);
                    stream_store(&vol[d * pl + o + x], (t1 + t2) + t3);
                }
            }
        }
    }
}

// seq/CombineCost.cpp:11-25.
__global__ void k_combine(float *__restrict__ c1, const float *__restrict__ c2, float l1, float l2, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const float a = 1.0f - expf(__fdiv_rn(-c1[i], l1));
        const float b = 1.0f - expf(__fdiv_rn(-c2[i], l2));
        c1[i] = a + b;
    }
}

}  // namespace rts

using namespace rts;

static rts_status check_rln(int W, int H, int D, int hr, int wr, float l1, float l2, float l3, int dch) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(hr >= 0 && wr >= 0 && hr <= NC_RMAX && wr <= NC_RMAX, "NCC radii (%d, %d) must be in [0, %d]", hr, wr,
              NC_RMAX);
    RTS_CHECK(2 * hr + 1 <= H && 2 * wr + 1 <= W, "NCC window larger than the image");
    RTS_CHECK(l1 > 0 && l2 > 0 && l3 > 0, "combine weights must be positive");
    RTS_CHECK(dch == 8 || dch == 16 || dch == 32, "disparities per block %d must be 8, 16 or 32", dch);
    RTS_CHECK(div_up(D, dch) <= 65535 && div_up(H, 4) <= 65535, "image too large");
    return RTS_OK;
}

template <bool PRE>
static void launch_rln_sweep(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                             const uint8_t *rank_tar, const float *log_ref, const float *log_tar, const uint4 *rec_ref,
                             const uint4 *rec_tar, float *vol, int W, int H, int D, float l1, float l2, float l3,
                             int dch, hipStream_t s) {
    dim3 grid(div_up(W, 64), div_up(H, NS_SEG), div_up(D, dch));
    if (dch == 8)
        k_cost_rln_sweep<3, 3, 8, PRE><<<grid, 256, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar,
                                                             rec_ref, rec_tar, vol, W, H, D, l1, l2, l3);
    else if (dch == 16)
        k_cost_rln_sweep<3, 3, 16, PRE><<<grid, 256, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar,
                                                              rec_ref, rec_tar, vol, W, H, D, l1, l2, l3);
    else
        k_cost_rln_sweep<3, 3, 32, PRE><<<grid, 256, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar,
                                                              rec_ref, rec_tar, vol, W, H, D, l1, l2, l3);
}

extern "C" {

static rts_status launch_ad(int kind, const void *ref, const void *tar, float *vol, int W, int H, int D, int dir,
                            int use_trunc, float trunc, int norm, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(dir == RTS_L2R || dir == RTS_R2L, "bad direction %d", dir);
    RTS_CHECK(H <= 65535, "height %d too large", H);
    dim3 grid(div_up(W, XB), H, div_up(D, DCH));
    if (kind == AD_GRAY)
        k_ad<AD_GRAY><<<grid, XB, 0, S(stream)>>>(ref, tar, vol, W, H, D, dir, 0, 0.f, 0);
    else if (kind == AD_FLOAT)
        k_ad<AD_FLOAT><<<grid, XB, 0, S(stream)>>>(ref, tar, vol, W, H, D, dir, 0, 0.f, 0);
    else if (W % 8 == 0 && W >= 1024 && W <= 7168 && (reinterpret_cast<uintptr_t>(vol) & 15) == 0)
        k_ad_color_h8<float><<<dim3(H, div_up(D, ADH_DCH)), 256, (size_t)(W + W / 8) * 8, S(stream)>>>(
            static_cast<const uint8_t *>(ref), static_cast<const uint8_t *>(tar), vol, W, H, D, dir, use_trunc, trunc,
            norm);
    else
        k_ad<AD_COLOR><<<grid, XB, 0, S(stream)>>>(ref, tar, vol, W, H, D, dir, use_trunc, trunc, norm);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_ad_gray(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int dir, void *stream) {
    return launch_ad(AD_GRAY, ref, tar, vol, W, H, D, dir, 0, 0.f, 0, stream);
}
rts_status rts_ad_float(const float *ref, const float *tar, float *vol, int W, int H, int D, int dir, void *stream) {
    return launch_ad(AD_FLOAT, ref, tar, vol, W, H, D, dir, 0, 0.f, 0, stream);
}
rts_status rts_ad_color(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int dir,
                        int use_trunc, float trunc, int norm, void *stream) {
    return launch_ad(AD_COLOR, ref, tar, vol, W, H, D, dir, use_trunc, trunc, norm, stream);
}

rts_status rts_ad_color_half(const uint8_t *ref, const uint8_t *tar, void *vol_half, int W, int H, int D, int dir,
                             int use_trunc, float trunc, int norm, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(dir == RTS_L2R || dir == RTS_R2L, "bad direction %d", dir);
    RTS_CHECK(H <= 65535, "height %d too large", H);
    // (rows of at least 1024 pixels: 128 threads of a block at work)
    if (W % 8 == 0 && W >= 1024 && W <= 7168 && (reinterpret_cast<uintptr_t>(vol_half) & 15) == 0) {
        dim3 grid(H, div_up(D, ADH_DCH));
        k_ad_color_h8<_Float16><<<grid, 256, (size_t)(W + W / 8) * 8, S(stream)>>>(
            ref, tar, static_cast<_Float16 *>(vol_half), W, H, D, dir, use_trunc, trunc, norm);
    } else {
        dim3 grid(div_up(W, XB), H, div_up(D, DCH));
        k_ad<AD_COLOR, _Float16><<<grid, XB, 0, S(stream)>>>(ref, tar, static_cast<_Float16 *>(vol_half), W, H, D,
                                                              dir, use_trunc, trunc, norm);
    }
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_census(const uint8_t *img, uint64_t *words, int W, int H, int hr, int wr, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size %dx%d", W, H);
    RTS_CHECK(hr >= 0 && wr >= 0 && hr < H && wr < W, "census radii %d,%d do not fit the image", hr, wr);
    dim3 grid(div_up(W, 64), div_up(H, 4)), block(64, 4);
    if (hr == 4 && wr == 3)
        k_census_w4<4, 3><<<dim3(div_up(W, 256), div_up(H, CW_TY)), block, 0, S(stream)>>>(img, words, W, H);
    else if (hr <= CT_R && wr <= CT_R)
        k_census_tile<<<grid, block, 0, S(stream)>>>(img, words, W, H, hr, wr);
    else
        k_census<<<grid, block, 0, S(stream)>>>(img, words, W, H, hr, wr);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_census_unpack(const uint64_t *words, uint8_t *bools, int W, int H, int N, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && N >= 1, "bad census size");
    const int P = W * H;
    k_census_unpack<<<div_up(P, 256), 256, 0, S(stream)>>>(words, bools, P, N);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_census_pack(const uint8_t *bools, uint64_t *words, int W, int H, int N, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1 && N >= 1, "bad census size");
    const int P = W * H;
    k_census_pack<<<div_up(P, 256), 256, 0, S(stream)>>>(bools, words, P, N);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_hamming(const uint64_t *cref, const uint64_t *ctar, float *vol, int W, int H, int D, int nwords,
                       void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(nwords >= 1, "nwords %d", nwords);
    RTS_CHECK(H <= 65535, "height %d too large", H);
    dim3 grid(div_up(W, XB), H, div_up(D, DCH));
    k_hamming<<<grid, XB, 0, S(stream)>>>(cref, ctar, vol, W, H, D, nwords);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_rank(const uint8_t *img, uint8_t *out, int W, int H, int m, void *stream) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size %dx%d", W, H);
    RTS_CHECK(m >= 0 && m < W && m < H, "rank radius %d does not fit the image", m);
    dim3 grid(div_up(W, 64), div_up(H, 4)), block(64, 4);
    k_rank<<<grid, block, 0, S(stream)>>>(img, out, W, H, m);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_convolve(const uint8_t *img, float *out, const float *mask_host, int W, int H, int hr, int wr,
                        void *stream) {
    RTS_CHECK(W >= 1 && H >= 1, "bad image size %dx%d", W, H);
    RTS_CHECK(hr >= 0 && wr >= 0 && hr < H && wr < W, "mask radii %d,%d do not fit the image", hr, wr);
    RTS_CHECK((2 * hr + 1) * (2 * wr + 1) <= 225, "mask of %dx%d taps exceeds 15x15", 2 * hr + 1, 2 * wr + 1);
    ConvMask m{};
    for (int i = 0; i < (2 * hr + 1) * (2 * wr + 1); i++) m.m[i] = mask_host[i];
    dim3 grid(div_up(W, 64), div_up(H, 4)), block(64, 4);
    k_convolve<<<grid, block, 0, S(stream)>>>(img, out, W, H, hr, wr, m);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_ncc(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int hr, int wr,
                   int flags, void *stream) {
    if (int e = check_dims(W, H, D)) return e;
    RTS_CHECK(hr >= 0 && wr >= 0 && hr < H && wr + D <= W, "NCC radii %d,%d do not fit the image", hr, wr);
    const int mode = flags & 0xff;
    const int keep = (flags & RTS_NCC_KEEP_FLAT) != 0;
    RTS_CHECK(mode == RTS_NCC_SEQ || mode == RTS_NCC_INTEGRAL, "bad NCC mode %d", mode);
    RTS_CHECK(div_up(D, DCH) <= 65535, "too many disparities");
    dim3 grid(div_up(W, 64), div_up(H, 4), div_up(D, DCH)), block(64, 4);
    if (mode == RTS_NCC_SEQ)
        k_ncc_seq<<<grid, block, 0, S(stream)>>>(ref, tar, vol, W, H, D, hr, wr, keep);
    else
        k_ncc_int<<<grid, block, 0, S(stream)>>>(ref, tar, vol, W, H, D, hr, wr, keep);
    RTS_LAUNCHED();
    return RTS_OK;
}
rts_status rts_cost_rank_log_ncc_dch(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                     const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol,
                                     int W, int H, int D, int hr, int wr, float l1, float l2, float l3, int dch,
                                     void *stream) {
    if (int e = check_rln(W, H, D, hr, wr, l1, l2, l3, dch)) return e;
    hipStream_t s = S(stream);
    if (hr == 3 && wr == 3) {   // the batched workload's window: row sweep
        launch_rln_sweep<false>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, nullptr, nullptr, vol, W, H,
                                D, l1, l2, l3, dch, s);
        RTS_LAUNCHED();
        return RTS_OK;
    }
    dim3 grid(div_up(W, 64), div_up(H, 4), div_up(D, dch)), block(64, 4);
    if (dch == 8)
        k_cost_rank_log_ncc<8><<<grid, block, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, vol, W,
                                                      H, D, hr, wr, l1, l2, l3);
    else if (dch == 16)
        k_cost_rank_log_ncc<16><<<grid, block, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, vol, W,
                                                       H, D, hr, wr, l1, l2, l3);
    else
        k_cost_rank_log_ncc<32><<<grid, block, 0, s>>>(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, vol, W,
                                                       H, D, hr, wr, l1, l2, l3);
    RTS_LAUNCHED();
    return RTS_OK;
}

size_t rts_cost_rank_log_ncc_ws_bytes(int W, int H) {
    if (W < 1 || H < 1) return 0;
    const size_t rec = ((size_t)W * (size_t)H * sizeof(uint4) + 255) & ~(size_t)255;
    return 2 * rec;
}

rts_status rts_cost_rank_log_ncc_ws(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                    const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol,
                                    int W, int H, int D, int hr, int wr, float l1, float l2, float l3, int dch,
                                    void *ws, size_t ws_bytes, void *stream) {
    if (hr != 3 || wr != 3 || ws == nullptr)
        return rts_cost_rank_log_ncc_dch(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, vol, W, H, D, hr,
                                         wr, l1, l2, l3, dch, stream);
    if (int e = check_rln(W, H, D, hr, wr, l1, l2, l3, dch)) return e;
    const size_t need = rts_cost_rank_log_ncc_ws_bytes(W, H);
    RTS_CHECK(ws_bytes >= need, "workspace of %zu bytes, %zu needed", ws_bytes, need);
    RTS_CHECK(((uintptr_t)ws & 15) == 0, "workspace must be 16-byte aligned");
    hipStream_t s = S(stream);
    uint4 *rec_ref = (uint4 *)ws, *rec_tar = (uint4 *)((uint8_t *)ws + need / 2);
    const dim3 rgrid(div_up(W, 64), div_up(H, 4));
    k_ncc_records<3, 3><<<rgrid, 256, 0, s>>>(gray_ref, rank_ref, log_ref, rec_ref, W, H);
    k_ncc_records<3, 3><<<rgrid, 256, 0, s>>>(gray_tar, rank_tar, log_tar, rec_tar, W, H);
    launch_rln_sweep<true>(gray_ref, gray_tar, nullptr, nullptr, nullptr, nullptr, rec_ref, rec_tar, vol, W, H, D, l1,
                           l2, l3, dch, s);
    RTS_LAUNCHED();
    return RTS_OK;
}

rts_status rts_cost_rank_log_ncc(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                 const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol, int W,
                                 int H, int D, int hr, int wr, float l1, float l2, float l3, void *stream) {
    return rts_cost_rank_log_ncc_dch(gray_ref, gray_tar, rank_ref, rank_tar, log_ref, log_tar, vol, W, H, D, hr, wr, l1,
                                     l2, l3, 16, stream);
}
rts_status rts_combine(float *c1, const float *c2, float l1, float l2, size_t n, void *stream) {
    if (!n) return RTS_OK;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 256 * 32);
    k_combine<<<blocks, 256, 0, S(stream)>>>(c1, c2, l1, l2, n);
    RTS_LAUNCHED();
    return RTS_OK;
}

}  // extern "C"
