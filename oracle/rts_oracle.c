/*
 * rts_oracle.c -- CPU oracle for the dense stereo pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the arithmetic of
 * the reference's sequential stage functions (haibo031031/rtstereo, the seq directory)
 * so that the HIP kernels can be checked against it.  Only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark may load it, and
 * only as a checker: the product path never calls it.
 *
 * Layouts are the reference's: images are row-major [H][W] (gray) or
 * [H][W][3] (colour); cost volumes are pixel-major cost[d + D*(y*W + x)];
 * census is bool[H][W][N]; arms are int[H][W][2].
 *
 * Every function keeps the reference's operation order and types where they
 * fix the rounding (float accumulation order, float vs int arithmetic, no
 * fused multiply-add: build with -ffp-contract=off).  Where the reference has
 * a quirk the function takes a flag: "literal" reproduces it, the default is
 * the intended behaviour the GPU path implements:
 *   - box aggregation in place vs out of place (seq/ConstBlockAggregation.cpp:27,31);
 *   - WTA carrying the previous pixel's disparity when d=0 wins (seq/WTA.cpp:12-15);
 *   - NCC leaving the cost unwritten when the denominator is 0
 *     (seq/NCCDisparity.cpp:70-71);
 *   - disparity refinement voting in place (seq/CrossBasedAggregation.cpp:171-212).
 * FP64 variants (suffix _f64) compute the same quantities in double precision
 * and serve as the accuracy reference for the float stages.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define IDX(y, x, W) ((size_t)(y) * (size_t)(W) + (size_t)(x))

/* One-step toroidal wrap, valid for |offset| < n (the reference's borders). */
static inline int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

/* ---------------------------------------------------------------- AD costs */

/* Gray AD, seq/PixelBasedCost.cpp:13-37.  dir=+1: tar at x-d (L2R),
 * dir=-1: tar at x+d (R2L).  |int difference| stored as float. */
void ora_ad_gray(const uint8_t *ref, const uint8_t *tar, float *cost, int D, int H, int W, int dir)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const int xt = wrap1(x - dir * d, W);
                const int diff = (int)ref[p] - (int)tar[IDX(y, xt, W)];
                cost[d + (size_t)D * p] = (float)(diff < 0 ? -diff : diff);
            }
        }
}

/* Float AD, seq/PixelBasedCost.cpp:39-69: one float subtraction, then |.|. */
void ora_ad_float(const float *ref, const float *tar, float *cost, int D, int H, int W, int dir)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const int xt = wrap1(x - dir * d, W);
                float r = ref[p] - tar[IDX(y, xt, W)];
                if (r < 0) r = -r;
                cost[d + (size_t)D * p] = r;
            }
        }
}

/* Colour AD, seq/PixelBasedCost.cpp:72-158.  The three channel |differences|
 * are summed as ints; with truncation the sum is replaced by trunc when it is
 * larger (compared as float, result float); with norm the value is multiplied
 * by the float 1/3.  use_trunc=0 is the overload without a trunc argument. */
void ora_ad_color(const uint8_t *ref, const uint8_t *tar, float *cost, int D, int H, int W, int dir,
                  int use_trunc, float trunc, int norm)
{
    const float third = (float)1 / (float)3;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const size_t q = IDX(y, wrap1(x - dir * d, W), W);
                int s = 0;
                for (int c = 0; c < 3; c++) {
                    const int t = (int)ref[3 * p + c] - (int)tar[3 * q + c];
                    s += t < 0 ? -t : t;
                }
                float v = use_trunc ? (((float)s > trunc) ? trunc : (float)s) : (float)s;
                if (norm) v = third * v;
                cost[d + (size_t)D * p] = v;
            }
        }
}

/* ------------------------------------------------------------ census */

/* seq/CensusTransform.cpp:11-37: entry k=(dy+hr)*(2wr+1)+(dx+wr) is 1 when the
 * centre is strictly brighter than the (wrapped) neighbour. */
void ora_census(const uint8_t *img, uint8_t *bits, int hr, int wr, int H, int W)
{
    const int N = (2 * hr + 1) * (2 * wr + 1);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            const uint8_t c = img[p];
            int k = 0;
            for (int dy = -hr; dy <= hr; dy++) {
                const int yy = wrap1(y + dy, H);
                for (int dx = -wr; dx <= wr; dx++, k++)
                    bits[k + (size_t)N * p] = c > img[IDX(yy, wrap1(x + dx, W), W)];
            }
        }
}

/* seq/CensusTransform.cpp:39-62: number of differing entries between the ref
 * descriptor at x and the tar descriptor at x-d (wrapped), as float. */
void ora_hamming(const uint8_t *bref, const uint8_t *btar, float *cost, int hr, int wr, int H, int W, int D)
{
    const int N = (2 * hr + 1) * (2 * wr + 1);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const size_t q = IDX(y, wrap1(x - d, W), W);
                int n = 0;
                for (int k = 0; k < N; k++) n += (bref[k + (size_t)N * p] != 0) != (btar[k + (size_t)N * q] != 0);
                cost[d + (size_t)D * p] = (float)n;
            }
        }
}

/* ------------------------------------------------------------ rank */

/* seq/RankFilter.cpp:11-34: 1 + number of window pixels darker than the
 * centre, radius m, wrapped, stored through an unsigned char cast. */
void ora_rank(const uint8_t *img, uint8_t *out, int m, int H, int W)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const uint8_t c = img[IDX(y, x, W)];
            int n = 1;
            for (int dy = -m; dy <= m; dy++) {
                const int yy = wrap1(y + dy, H);
                for (int dx = -m; dx <= m; dx++)
                    n += img[IDX(yy, wrap1(x + dx, W), W)] < c;
            }
            out[IDX(y, x, W)] = (uint8_t)n;
        }
}

/* ------------------------------------------------------------ LoG */

/* seq/ConvolveImage.cpp:11-34: float correlation in row-major window order
 * (product rounded to float, then added), |.|, clamped to 255. */
void ora_convolve(const uint8_t *img, float *out, const float *mask, int H, int W, int hr, int wr)
{
    const int mw = 2 * wr + 1;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float acc = 0.0f;
            for (int dy = -hr; dy <= hr; dy++) {
                const int yy = wrap1(y + dy, H);
                for (int dx = -wr; dx <= wr; dx++) {
                    const float prod = (float)img[IDX(yy, wrap1(x + dx, W), W)] * mask[(dy + hr) * mw + (dx + wr)];
                    acc += prod;
                }
            }
            acc = acc > 0 ? acc : -acc;
            out[IDX(y, x, W)] = acc > 255 ? 255.0f : acc;
        }
}

void ora_convolve_f64(const uint8_t *img, double *out, const float *mask, int H, int W, int hr, int wr)
{
    const int mw = 2 * wr + 1;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            double acc = 0.0;
            for (int dy = -hr; dy <= hr; dy++)
                for (int dx = -wr; dx <= wr; dx++)
                    acc += (double)img[IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)] * (double)mask[(dy + hr) * mw + (dx + wr)];
            acc = fabs(acc);
            out[IDX(y, x, W)] = acc > 255 ? 255.0 : acc;
        }
}

/* ------------------------------------------------------------ NCC */

/* seq/NCCDisparity.cpp:11-73.  Means are float sums divided by the window
 * area; a' = ref - meanRef, b' = tar - meanTar; nom = sum a'b',
 * den = sqrt(sum a'a'b'b') (the reference's non-standard denominator).
 * den == 0: literal=1 leaves the cost untouched, literal=0 writes 0. */
void ora_ncc(const uint8_t *ref, const uint8_t *tar, float *cost, int D, int H, int W, int hr, int wr, int literal)
{
    const int area = (2 * hr + 1) * (2 * wr + 1);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            float mr = 0.0f;
            for (int dy = -hr; dy <= hr; dy++)
                for (int dx = -wr; dx <= wr; dx++)
                    mr += (float)ref[IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)];
            mr = mr / (float)area;
            for (int d = 0; d < D; d++) {
                float mt = 0.0f;
                for (int dy = -hr; dy <= hr; dy++)
                    for (int dx = -wr; dx <= wr; dx++)
                        mt += (float)tar[IDX(wrap1(y + dy, H), wrap1(x + dx - d, W), W)];
                mt = mt / (float)area;
                float nom = 0.0f, den = 0.0f;
                for (int dy = -hr; dy <= hr; dy++) {
                    const int yy = wrap1(y + dy, H);
                    for (int dx = -wr; dx <= wr; dx++) {
                        const float a = (float)ref[IDX(yy, wrap1(x + dx, W), W)] - mr;
                        const float b = (float)tar[IDX(yy, wrap1(x + dx - d, W), W)] - mt;
                        const float ab = a * b;
                        nom += ab;
                        const float aa = a * a;
                        const float aab = aa * b;
                        const float aabb = aab * b;
                        den += aabb;
                    }
                }
                den = sqrtf(den);
                if (den != 0.0f)
                    cost[d + (size_t)D * p] = nom / den;
                else if (!literal)
                    cost[d + (size_t)D * p] = 0.0f;
            }
        }
}

/* The same quantity in double precision from exact window sums; 0 where the
 * denominator is 0. */
void ora_ncc_f64(const uint8_t *ref, const uint8_t *tar, double *cost, int D, int H, int W, int hr, int wr)
{
    const int area = (2 * hr + 1) * (2 * wr + 1);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            double mr = 0.0;
            for (int dy = -hr; dy <= hr; dy++)
                for (int dx = -wr; dx <= wr; dx++)
                    mr += ref[IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)];
            mr /= area;
            for (int d = 0; d < D; d++) {
                double mt = 0.0;
                for (int dy = -hr; dy <= hr; dy++)
                    for (int dx = -wr; dx <= wr; dx++)
                        mt += tar[IDX(wrap1(y + dy, H), wrap1(x + dx - d, W), W)];
                /* The denominator is exactly 0 iff every tap has a centred ref
                 * or tar value of 0: decide that in integers, since the
                 * rounded double sum can leave a tiny non-zero residue. */
                const long sa = (long)(mr * area + 0.5), sb = (long)(mt + 0.5);
                int flat = 1;
                for (int dy = -hr; dy <= hr && flat; dy++)
                    for (int dx = -wr; dx <= wr; dx++) {
                        const long a = (long)area * ref[IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)] - sa;
                        const long b = (long)area * tar[IDX(wrap1(y + dy, H), wrap1(x + dx - d, W), W)] - sb;
                        if (a != 0 && b != 0) { flat = 0; break; }
                    }
                mt /= area;
                if (flat) {
                    cost[d + (size_t)D * p] = 0.0;
                    continue;
                }
                double nom = 0.0, den = 0.0;
                for (int dy = -hr; dy <= hr; dy++)
                    for (int dx = -wr; dx <= wr; dx++) {
                        const double a = ref[IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)] - mr;
                        const double b = tar[IDX(wrap1(y + dy, H), wrap1(x + dx - d, W), W)] - mt;
                        nom += a * b;
                        den += a * a * b * b;
                    }
                den = sqrt(den);
                cost[d + (size_t)D * p] = den != 0.0 ? nom / den : 0.0;
            }
        }
}

/* ------------------------------------------------------------ combine */

/* seq/CombineCost.cpp:11-25: c1 <- (1 - expf(-c1/l1)) + (1 - expf(-c2/l2)). */
void ora_combine(float *c1, const float *c2, float l1, float l2, int H, int W, int D)
{
    const size_t n = (size_t)H * W * D;
    for (size_t i = 0; i < n; i++) {
        const float a = 1 - expf(-c1[i] / l1);
        const float b = 1 - expf(-c2[i] / l2);
        c1[i] = a + b;
    }
}

/* Three-term form of the batched workload (the seq/CombineCost.cpp:11-25
 * term for each cost; the NCC similarity enters as 1 - ncc):
 *   c1 = ((1-expf(-c1/l1)) + (1-expf(-c2/l2))) + (1-expf(-(1-ncc)/l3)). */
void ora_combine3(float *c1, const float *c2, const float *ncc, float l1, float l2, float l3, size_t n)
{
    for (size_t i = 0; i < n; i++) {
        const float a = 1 - expf(-c1[i] / l1);
        const float b = 1 - expf(-c2[i] / l2);
        const float c = 1 - expf(-(1 - ncc[i]) / l3);
        c1[i] = (a + b) + c;
    }
}

void ora_combine_f64(const float *c1, const float *c2, double l1, double l2, double *out, size_t n)
{
    for (size_t i = 0; i < n; i++)
        out[i] = (1.0 - exp(-(double)c1[i] / l1)) + (1.0 - exp(-(double)c2[i] / l2));
}

/* ------------------------------------------------------------ box */

/* seq/ConstBlockAggregation.cpp:11-34: window mean with wrapped borders,
 * float sum in row-major window order, then sum/(float)area.
 * literal=1: in place on `cost` (later pixels read already aggregated
 * neighbours, as the reference does); literal=0: out of place into `out`. */
void ora_box(float *cost, float *out, int H, int W, int D, int r, int literal)
{
    const int area = (2 * r + 1) * (2 * r + 1);
    float *dst = literal ? cost : out;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
            for (int d = 0; d < D; d++) {
                float acc = 0.0f;
                for (int dy = -r; dy <= r; dy++) {
                    const int yy = wrap1(y + dy, H);
                    for (int dx = -r; dx <= r; dx++)
                        acc += cost[d + (size_t)D * IDX(yy, wrap1(x + dx, W), W)];
                }
                dst[d + (size_t)D * IDX(y, x, W)] = acc / (float)area;
            }
}

/* ------------------------------------------------------------ adaptive weight */

/* seq/AdaptiveWeightAggregation.cpp:11-18: exp(-sqrt(dy^2+dx^2)/gamma_p). */
void ora_prox_weight(float *prox, float gamma_p, int r)
{
    int k = 0;
    for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++, k++)
            prox[k] = expf(-sqrtf((float)(dy * dy + dx * dx)) / gamma_p);
}

/* seq/AdaptiveWeightAggregation.cpp:19-24: Lab Euclidean distance in float. */
static inline float lab_dist(int l0, int l1, int a0, int a1, int b0, int b1)
{
    const float e0 = (float)l0 - (float)l1, e1 = (float)a0 - (float)a1, e2 = (float)b0 - (float)b1;
    const float s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    return sqrtf(s0 + s1 + s2);
}

/* seq/AdaptiveWeightAggregation.cpp:33-71: w[p][k] = prox[k]*expf(-dist/gamma_c),
 * 0 where the window position falls outside the image (no wrap here). */
void ora_compute_weights(float *w, const float *prox, const uint8_t *lab, float gamma_c, int H, int W, int r)
{
    const int e = 2 * r + 1, A = e * e;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            const uint8_t *c = lab + 3 * p;
            int k = 0;
            for (int dy = -r; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++, k++) {
                    const int yy = y + dy, xx = x + dx;
                    float v = 0.0f;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const uint8_t *t = lab + 3 * IDX(yy, xx, W);
                        const float dist = lab_dist(c[0], t[0], c[1], t[1], c[2], t[2]);
                        v = prox[k] * expf(-dist / gamma_c);
                    }
                    w[k + (size_t)A * p] = v;
                }
        }
}

/* seq/AdaptiveWeightAggregation.cpp:73-111: for each (p, d) with target
 * centre q = (x-d wrapped, y): sum over window offsets k of
 * w_ref[p][k]*w_tar[q][k] (float product), score += cost(p+k wrapped, d)*weight,
 * in row-major window order; result score/sum_weight. */
void ora_aw_l2r(const float *wref, const float *wtar, const float *cost, float *out, int H, int W, int D, int r)
{
    const int e = 2 * r + 1, A = e * e;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const size_t q = IDX(y, wrap1(x - d, W), W);
                float sw = 0.0f, sc = 0.0f;
                int k = 0;
                for (int dy = -r; dy <= r; dy++) {
                    const int yy = wrap1(y + dy, H);
                    for (int dx = -r; dx <= r; dx++, k++) {
                        const float wgt = wref[k + (size_t)A * p] * wtar[k + (size_t)A * q];
                        sw += wgt;
                        const float t = cost[d + (size_t)D * IDX(yy, wrap1(x + dx, W), W)] * wgt;
                        sc += t;
                    }
                }
                out[d + (size_t)D * p] = sc / sw;
            }
        }
}

/* FP64 adaptive-weight aggregation for a list of pixels (all d), computing the
 * weights on the fly in double: w = exp(-sqrt(dy^2+dx^2)/gp) * exp(-|Lab diff|/gc),
 * 0 outside the image.  out[i*D + d] for pixel i = (ys[i], xs[i]). */
void ora_aw_l2r_f64(const uint8_t *labL, const uint8_t *labR, const float *cost, const int *ys, const int *xs,
                    int npix, double *out, int H, int W, int D, int r, double gamma_c, double gamma_p)
{
    const int e = 2 * r + 1;
    double *prox = (double *)malloc(sizeof(double) * e * e);
    for (int dy = -r, k = 0; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++, k++) prox[k] = exp(-sqrt((double)(dy * dy + dx * dx)) / gamma_p);
    double *wl = (double *)malloc(sizeof(double) * e * e);
    for (int i = 0; i < npix; i++) {
        const int y = ys[i], x = xs[i];
        const uint8_t *c = labL + 3 * IDX(y, x, W);
        for (int dy = -r, k = 0; dy <= r; dy++)
            for (int dx = -r; dx <= r; dx++, k++) {
                const int yy = y + dy, xx = x + dx;
                double v = 0.0;
                if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                    const uint8_t *t = labL + 3 * IDX(yy, xx, W);
                    const double e0 = (double)c[0] - t[0], e1 = (double)c[1] - t[1], e2 = (double)c[2] - t[2];
                    v = prox[k] * exp(-sqrt(e0 * e0 + e1 * e1 + e2 * e2) / gamma_c);
                }
                wl[k] = v;
            }
        for (int d = 0; d < D; d++) {
            const int xq = wrap1(x - d, W);
            const uint8_t *cq = labR + 3 * IDX(y, xq, W);
            double sw = 0.0, sc = 0.0;
            for (int dy = -r, k = 0; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++, k++) {
                    const int yy = y + dy, xx = xq + dx;
                    double wt = 0.0;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const uint8_t *t = labR + 3 * IDX(yy, xx, W);
                        const double e0 = (double)cq[0] - t[0], e1 = (double)cq[1] - t[1], e2 = (double)cq[2] - t[2];
                        wt = prox[k] * exp(-sqrt(e0 * e0 + e1 * e1 + e2 * e2) / gamma_c);
                    }
                    const double wgt = wl[k] * wt;
                    sw += wgt;
                    sc += wgt * (double)cost[d + (size_t)D * IDX(wrap1(y + dy, H), wrap1(x + dx, W), W)];
                }
            out[(size_t)i * D + d] = sc / sw;
        }
    }
    free(wl);
    free(prox);
}

/* ora_aw_l2r_f64 for whole rows: out[i][x][d] for y = rows[i], every x and
 * d, with each row's target weights computed once.  Rows run one after the
 * other and pixels in parallel (OpenMP when built with it). */
void ora_aw_rows_f64(const uint8_t *labL, const uint8_t *labR, const float *cost, const int *rows, int nrows,
                     double *out, int H, int W, int D, int r, double gamma_c, double gamma_p)
{
    const int e = 2 * r + 1, A = e * e;
    double *prox = (double *)malloc(sizeof(double) * A);
    for (int dy = -r, k = 0; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++, k++) prox[k] = exp(-sqrt((double)(dy * dy + dx * dx)) / gamma_p);
    double *wt = (double *)malloc(sizeof(double) * (size_t)A * W);
    for (int i = 0; i < nrows; i++) {
        const int y = rows[i];
#pragma omp parallel for schedule(static)
        for (int x = 0; x < W; x++) {
            const uint8_t *c = labR + 3 * IDX(y, x, W);
            for (int dy = -r, k = 0; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++, k++) {
                    const int yy = y + dy, xx = x + dx;
                    double v = 0.0;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const uint8_t *t = labR + 3 * IDX(yy, xx, W);
                        const double e0 = (double)c[0] - t[0], e1 = (double)c[1] - t[1], e2 = (double)c[2] - t[2];
                        v = prox[k] * exp(-sqrt(e0 * e0 + e1 * e1 + e2 * e2) / gamma_c);
                    }
                    wt[(size_t)x * A + k] = v;
                }
        }
#pragma omp parallel for schedule(dynamic, 8)
        for (int x = 0; x < W; x++) {
            double wl[33 * 33];
            const uint8_t *c = labL + 3 * IDX(y, x, W);
            for (int dy = -r, k = 0; dy <= r; dy++)
                for (int dx = -r; dx <= r; dx++, k++) {
                    const int yy = y + dy, xx = x + dx;
                    double v = 0.0;
                    if (yy >= 0 && yy < H && xx >= 0 && xx < W) {
                        const uint8_t *t = labL + 3 * IDX(yy, xx, W);
                        const double e0 = (double)c[0] - t[0], e1 = (double)c[1] - t[1], e2 = (double)c[2] - t[2];
                        v = prox[k] * exp(-sqrt(e0 * e0 + e1 * e1 + e2 * e2) / gamma_c);
                    }
                    wl[k] = v;
                }
            for (int d = 0; d < D; d++) {
                const double *wq = wt + (size_t)wrap1(x - d, W) * A;
                double sw = 0.0, sc = 0.0;
                for (int dy = -r, k = 0; dy <= r; dy++) {
                    const float *crow = cost + (size_t)D * IDX(wrap1(y + dy, H), 0, W) + d;
                    for (int dx = -r; dx <= r; dx++, k++) {
                        const double wgt = wl[k] * wq[k];
                        sw += wgt;
                        sc += wgt * (double)crow[(size_t)D * wrap1(x + dx, W)];
                    }
                }
                out[((size_t)i * W + x) * D + d] = sc / sw;
            }
        }
    }
    free(wt);
    free(prox);
}

/* ------------------------------------------------------------ cross-based */

/* Maximum channel |difference| between two colour pixels. */
static inline int chan_delta(const uint8_t *a, const uint8_t *b)
{
    int m = 0;
    for (int c = 0; c < 3; c++) {
        const int t = abs((int)a[c] - (int)b[c]);
        if (t > m) m = t;
    }
    return m;
}

/* One arm, seq/CrossBasedAggregation.cpp:12-100 (calc_x_range/calc_y_range):
 * step outward while the previous step's colour delta is below tao, fewer
 * than L steps were taken and the next position is inside the image; the
 * step that first reaches tao is counted.  (step_y, step_x) is the direction. */
static int arm_len(const uint8_t *img, int H, int W, int y, int x, int sy, int sx, int L, int tao)
{
    const uint8_t *c = img + 3 * IDX(y, x, W);
    int n = 0, delta = 0;
    while (delta < tao && n < L) {
        const int yy = y + sy * (n + 1), xx = x + sx * (n + 1);
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) break;
        n++;
        delta = chan_delta(c, img + 3 * IDX(yy, xx, W));
    }
    /* The reference bumps a zero arm to 1 when the first neighbour is inside
     * the image (seq/CrossBasedAggregation.cpp:34-35, 52-53); that only
     * matters for L < 1. */
    if (n == 0) {
        const int yy = y + sy, xx = x + sx;
        if (yy >= 0 && yy < H && xx >= 0 && xx < W) n = 1;
    }
    return n;
}

/* seq/CrossBasedAggregation.cpp:102-115: horz[2p]=left, horz[2p+1]=right,
 * vert[2p]=top, vert[2p+1]=bottom. */
void ora_find_cross(const uint8_t *img, int *horz, int *vert, int H, int W, int L, int tao)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            vert[2 * p] = arm_len(img, H, W, y, x, -1, 0, L, tao);
            vert[2 * p + 1] = arm_len(img, H, W, y, x, 1, 0, L, tao);
            horz[2 * p] = arm_len(img, H, W, y, x, 0, -1, L, tao);
            horz[2 * p + 1] = arm_len(img, H, W, y, x, 0, 1, L, tao);
        }
}

/* seq/CrossBasedAggregation.cpp:117-125: serial float prefix along x, in place. */
void ora_hprefix(float *cost, int H, int W, int D)
{
    for (int y = 0; y < H; y++)
        for (int x = 1; x < W; x++)
            for (int d = 0; d < D; d++)
                cost[d + (size_t)D * IDX(y, x, W)] += cost[d + (size_t)D * IDX(y, x - 1, W)];
}

static inline int imin(int a, int b) { return a < b ? a : b; }

/* seq/CrossBasedAggregation.cpp:127-169.  P is the row prefix volume.  For
 * (x,y,d), tx = x-d wrapped; rows y-t..y+b with t,b = min of ref arm at
 * (x,y) and tar arm at (tx,y); per row l,r = min of ref arm at (x,row) and tar
 * arm at (tx,row); segment P[x+r] - P[x-l-1] with x-l-1 clamped to 0 (which
 * drops column 0 when the arm reaches it); count l+r+1; float sums in row
 * order, then sum / (float)count. */
void ora_cross_v(const float *P, float *out, const int *href, const int *htar, const int *vref, const int *vtar,
                 int H, int W, int D)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const int tx = wrap1(x - d, W);
                const size_t q = IDX(y, tx, W);
                const int t = imin(vref[2 * p], vtar[2 * q]);
                const int b = imin(vref[2 * p + 1], vtar[2 * q + 1]);
                float sum = 0.0f;
                int cnt = 0;
                for (int yy = y - t; yy <= y + b; yy++) {
                    if (yy < 0 || yy >= H) continue;
                    const size_t pr = IDX(yy, x, W), pt = IDX(yy, tx, W);
                    const int l = imin(href[2 * pr], htar[2 * pt]);
                    const int r = imin(href[2 * pr + 1], htar[2 * pt + 1]);
                    int xl = x - l - 1, xr = x + r;
                    if (xl < 0) xl = 0;
                    else if (xr >= W) xr = W - 1;
                    const float seg = P[d + (size_t)D * IDX(yy, xr, W)] - P[d + (size_t)D * IDX(yy, xl, W)];
                    sum += seg;
                    cnt += l + r + 1;
                }
                out[d + (size_t)D * p] = sum / (float)cnt;
            }
        }
}

/* The same support region summed exactly in double from the raw costs
 * (including the column-0 drop), then divided by the count. */
void ora_cross_f64(const float *cost, double *out, const int *href, const int *htar, const int *vref, const int *vtar,
                   int H, int W, int D)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            for (int d = 0; d < D; d++) {
                const int tx = wrap1(x - d, W);
                const size_t q = IDX(y, tx, W);
                const int t = imin(vref[2 * p], vtar[2 * q]);
                const int b = imin(vref[2 * p + 1], vtar[2 * q + 1]);
                double sum = 0.0;
                long cnt = 0;
                for (int yy = y - t; yy <= y + b; yy++) {
                    const size_t pr = IDX(yy, x, W), pt = IDX(yy, tx, W);
                    const int l = imin(href[2 * pr], htar[2 * pt]);
                    const int r = imin(href[2 * pr + 1], htar[2 * pt + 1]);
                    const int x0 = (x - l - 1 < 0) ? 1 : x - l;
                    for (int xx = x0; xx <= x + r; xx++) sum += cost[d + (size_t)D * IDX(yy, xx, W)];
                    cnt += l + r + 1;
                }
                out[d + (size_t)D * p] = sum / (double)cnt;
            }
        }
}

/* seq/CrossBasedAggregation.cpp:171-212: histogram vote over the cross
 * support (rows y-t .. y+b-1, columns x-l .. x+r clamped), most frequent
 * disparity with the lowest winning ties.  literal=1 votes in place (later
 * pixels see refined neighbours, as the reference does); literal=0 reads
 * `disp` and writes `out`. */
void ora_disp_refine(uint8_t *disp, uint8_t *out, const int *horz, const int *vert, int H, int W, int D, int literal)
{
    int *hist = (int *)malloc(sizeof(int) * (D > 0 ? D : 1));
    uint8_t *dst = literal ? disp : out;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const size_t p = IDX(y, x, W);
            memset(hist, 0, sizeof(int) * D);
            for (int yy = y - vert[2 * p]; yy < y + vert[2 * p + 1]; yy++) {
                if (yy < 0 || yy >= H) continue;
                const size_t pr = IDX(yy, x, W);
                int xl = x - horz[2 * pr], xr = x + horz[2 * pr + 1];
                if (xl < 0) xl = 0;
                else if (xr >= W) xr = W - 1;
                for (int xx = xl; xx <= xr; xx++) hist[disp[IDX(yy, xx, W)]]++;
            }
            int best = 0, bv = hist[0];
            for (int i = 1; i < D; i++)
                if (hist[i] > bv) { bv = hist[i]; best = i; }
            dst[p] = (uint8_t)best;
        }
    free(hist);
}

/* ------------------------------------------------------------ WTA / cross-check */

/* seq/WTA.cpp:10-39.  maximize=0: argmin (WTA), 1: argmax (WTA_NCC); strict
 * comparison from d=1 against the d=0 value, so the lowest d wins ties.
 * literal=1: when no d>0 beats d=0 the pixel keeps the previous pixel's
 * disparity (the reference never resets it); literal=0 writes 0 then. */
void ora_wta(const float *cost, uint8_t *disp, int H, int W, int D, int maximize, int literal)
{
    uint8_t carry = 0;
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const float *c = cost + (size_t)D * IDX(y, x, W);
            float best = c[0];
            uint8_t bd = literal ? carry : 0;
            for (int d = 1; d < D; d++)
                if (maximize ? (c[d] > best) : (c[d] < best)) { best = c[d]; bd = (uint8_t)d; }
            disp[IDX(y, x, W)] = bd;
            carry = bd;
        }
}

/* Right-view disparity read off the left-view volume's diagonal:
 * D_R(x) = argmin_d cost((x+d) wrapped, d), lowest d on ties. */
void ora_wta_r2l_diag(const float *cost, uint8_t *disp, int H, int W, int D, int maximize)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            float best = cost[(size_t)D * IDX(y, x, W)];
            uint8_t bd = 0;
            for (int d = 1; d < D; d++) {
                const float v = cost[d + (size_t)D * IDX(y, wrap1(x + d, W), W)];
                if (maximize ? (v > best) : (v < best)) { best = v; bd = (uint8_t)d; }
            }
            disp[IDX(y, x, W)] = bd;
        }
}

/* seq/CrossCheck.cpp:10-23 with a tolerance: 1 where
 * |D_L(x) - D_R(x - D_L(x) wrapped)| <= tol (tol 0 is the reference). */
void ora_cross_check(const uint8_t *dl, const uint8_t *dr, uint8_t *mask, int H, int W, int tol)
{
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int v = dl[IDX(y, x, W)];
            const int t = dr[IDX(y, wrap1(x - v, W), W)];
            mask[IDX(y, x, W)] = abs(v - t) <= tol;
        }
}
