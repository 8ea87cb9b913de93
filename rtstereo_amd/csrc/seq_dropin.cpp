// Drop-in implementation of the reference's seq/ stage functions on top of
// librtstereo's C ABI.  Each call uploads the caller's host buffers,
// converts pixel-major volumes to the device's disparity-major layout, runs
// the HIP kernels, converts back and downloads, all on one stream per host
// thread with grow-only device buffers cached per thread.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "rtstereo.h"
#include "rtstereo_seq.h"

#include <algorithm>

namespace {

thread_local std::string g_seq_err;

struct Ctx {
    void *stream = nullptr;
    std::vector<std::pair<void *, size_t>> slots;
    bool ok = false;
    Ctx() { ok = rts_stream_create(&stream) == RTS_OK; }
    ~Ctx() {
        for (auto &s : slots) rts_free(s.first);
        if (stream) rts_stream_destroy(stream);
    }
    // Device buffer `i` of at least `bytes` (contents not preserved on growth).
    void *buf(int i, size_t bytes) {
        if ((int)slots.size() <= i) slots.resize(i + 1, {nullptr, 0});
        auto &s = slots[i];
        if (s.second < bytes) {
            rts_free(s.first);
            s.first = nullptr;
            s.second = 0;
            if (rts_malloc(&s.first, bytes) != RTS_OK) return nullptr;
            s.second = bytes;
        }
        return s.first;
    }
};

Ctx &ctx() {
    thread_local Ctx c;
    return c;
}

// Reports a failure: abort with the message unless RTS_SEQ_NOABORT is set.
bool failed(int st, const char *fn) {
    if (st == RTS_OK) return false;
    g_seq_err = std::string(fn) + ": " + rts_last_error();
    const char *env = getenv("RTS_SEQ_NOABORT");
    if (!(env && env[0] == '1')) {
        fprintf(stderr, "rtstereo: %s\n", g_seq_err.c_str());
        abort();
    }
    return true;
}

#define STEP(call)                          \
    do {                                    \
        if (failed((call), __func__)) return; \
    } while (0)

#define NEED(ptr)                                                                             \
    do {                                                                                      \
        if (!(ptr)) {                                                                         \
            failed(RTS_ERR_NOMEM, __func__);                                                  \
            return;                                                                           \
        }                                                                                     \
    } while (0)

struct Call {
    Ctx &c;
    bool ok;
    Call() : c(ctx()), ok(c.ok) { g_seq_err.clear(); }
    void *dev(int slot, size_t bytes) { return c.buf(slot, bytes); }
    void *s() { return c.stream; }
};

size_t P(int rdim, int cdim) { return (size_t)rdim * (size_t)cdim; }

// Uploads a pixel-major host volume into disparity-major device memory.
int vol_up(Call &k, const float *host, float *dm, int W, int H, int D) {
    float *stage = (float *)k.dev(15, P(H, W) * D * sizeof(float));
    if (!stage) return RTS_ERR_NOMEM;
    if (int e = rts_memcpy_h2d(stage, host, P(H, W) * D * sizeof(float), k.s())) return e;
    return rts_vol_to_dm(stage, dm, W, H, D, k.s());
}

// Downloads a disparity-major device volume into a pixel-major host volume.
int vol_down(Call &k, const float *dm, float *host, int W, int H, int D) {
    float *stage = (float *)k.dev(15, P(H, W) * D * sizeof(float));
    if (!stage) return RTS_ERR_NOMEM;
    if (int e = rts_vol_to_pm(dm, stage, W, H, D, k.s())) return e;
    if (int e = rts_memcpy_d2h(host, stage, P(H, W) * D * sizeof(float), k.s())) return e;
    return rts_stream_sync(k.s());
}

// AD family: upload two images of `px_bytes` per pixel, build, download.
void ad_call(const char *fn, int kind, const void *ref, const void *tar, float *cost, int D, int H, int W, int dir,
             int use_trunc, float trunc, int norm) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, fn)) return;
    const size_t pxb = kind == 0 ? 1 : (kind == 1 ? 4 : 3);
    void *a = k.dev(0, P(H, W) * pxb), *b = k.dev(1, P(H, W) * pxb);
    float *vol = (float *)k.dev(2, P(H, W) * D * sizeof(float));
    if (!a || !b || !vol) {
        failed(RTS_ERR_NOMEM, fn);
        return;
    }
    if (failed(rts_memcpy_h2d(a, ref, P(H, W) * pxb, k.s()), fn)) return;
    if (failed(rts_memcpy_h2d(b, tar, P(H, W) * pxb, k.s()), fn)) return;
    int st;
    if (kind == 0)
        st = rts_ad_gray((const uint8_t *)a, (const uint8_t *)b, vol, W, H, D, dir, k.s());
    else if (kind == 1)
        st = rts_ad_float((const float *)a, (const float *)b, vol, W, H, D, dir, k.s());
    else
        st = rts_ad_color((const uint8_t *)a, (const uint8_t *)b, vol, W, H, D, dir, use_trunc, trunc, norm, k.s());
    if (failed(st, fn)) return;
    failed(vol_down(k, vol, cost, W, H, D), fn);
}

// Arms from the reference's int pairs, validated on the host (the packed
// device layout holds 0..255 per arm).
int arms_up(Call &k, const int *horz, const int *vert, uint32_t *dst, int W, int H, int slot) {
    const size_t n = P(H, W);
    for (size_t i = 0; i < 2 * n; i++)
        if (horz[i] < 0 || horz[i] > 255 || vert[i] < 0 || vert[i] > 255) return RTS_ERR_INVALID;
    int *hd = (int *)k.dev(slot, n * 2 * sizeof(int)), *vd = (int *)k.dev(slot + 1, n * 2 * sizeof(int));
    if (!hd || !vd) return RTS_ERR_NOMEM;
    if (int e = rts_memcpy_h2d(hd, horz, n * 2 * sizeof(int), k.s())) return e;
    if (int e = rts_memcpy_h2d(vd, vert, n * 2 * sizeof(int), k.s())) return e;
    if (int e = rts_arms_pack(hd, vd, dst, W, H, k.s())) return e;
    // hd/vd are reused by the next arms_up call only after this stream's
    // work, which is ordered on the same stream.
    return RTS_OK;
}

}  // namespace

extern "C" const char *rts_seq_last_error(void) { return g_seq_err.c_str(); }

// ---------------------------------------------------------------- AD (seq/PixelBasedCost.cpp)

void pixelBasedL2R(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                   const int rdim, const int cdim) {
    ad_call(__func__, 0, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_L2R, 0, 0.f, 0);
}
void pixelBasedR2L(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                   const int rdim, const int cdim) {
    ad_call(__func__, 0, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_R2L, 0, 0.f, 0);
}
void pixelBasedCostL2R_Float(float *ref_image, float *tar_image, float *cost, const int dispRange, const int rdim,
                             const int cdim) {
    ad_call(__func__, 1, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_L2R, 0, 0.f, 0);
}
void pixelBasedCostR2L_Float(float *ref_image, float *tar_image, float *cost, const int dispRange, const int rdim,
                             const int cdim) {
    ad_call(__func__, 1, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_R2L, 0, 0.f, 0);
}
void pixelBasedCostL2R_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, int normOrNot) {
    ad_call(__func__, 2, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_L2R, 0, 0.f, normOrNot != 0);
}
void pixelBasedCostL2R_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, float trunc, int normOrNot) {
    ad_call(__func__, 2, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_L2R, 1, trunc, normOrNot != 0);
}
void pixelBasedR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                         const int rdim, const int cdim, int normOrNot) {
    ad_call(__func__, 2, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_R2L, 0, 0.f, normOrNot != 0);
}
void pixelBasedR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                         const int rdim, const int cdim, float trunc, int normOrNot) {
    ad_call(__func__, 2, ref_image, tar_image, cost, dispRange, rdim, cdim, RTS_R2L, 1, trunc, normOrNot != 0);
}
void pixelBasedCostR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, int normOrNot) {
    pixelBasedR2L_Color(ref_image, tar_image, cost, dispRange, rdim, cdim, normOrNot);
}
void pixelBasedCostR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, float trunc, int normOrNot) {
    pixelBasedR2L_Color(ref_image, tar_image, cost, dispRange, rdim, cdim, trunc, normOrNot);
}

// ---------------------------------------------------------------- census (seq/CensusTransform.cpp)

void calcCensusTrans(unsigned char *img, bool *censusTrans, const int maskHrad, const int maskWrad, const int rdim,
                     const int cdim) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const int N = (2 * maskHrad + 1) * (2 * maskWrad + 1), nw = (N + 63) / 64;
    uint8_t *im = (uint8_t *)k.dev(0, P(rdim, cdim));
    uint64_t *words = (uint64_t *)k.dev(1, P(rdim, cdim) * nw * 8);
    uint8_t *bools = (uint8_t *)k.dev(2, P(rdim, cdim) * N);
    NEED(im && words && bools);
    STEP(rts_memcpy_h2d(im, img, P(rdim, cdim), k.s()));
    STEP(rts_census(im, words, cdim, rdim, maskHrad, maskWrad, k.s()));
    STEP(rts_census_unpack(words, bools, cdim, rdim, N, k.s()));
    STEP(rts_memcpy_d2h(censusTrans, bools, P(rdim, cdim) * N, k.s()));
    STEP(rts_stream_sync(k.s()));
}

void calcHammingDist(bool *censusTrnsRef, bool *censusTrnsTar, float *cost, const int maskHrad, const int maskWrad,
                     int rdim, int cdim, int dispRange) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const int N = (2 * maskHrad + 1) * (2 * maskWrad + 1), nw = (N + 63) / 64;
    uint8_t *b0 = (uint8_t *)k.dev(0, P(rdim, cdim) * N), *b1 = (uint8_t *)k.dev(1, P(rdim, cdim) * N);
    uint64_t *w0 = (uint64_t *)k.dev(3, P(rdim, cdim) * nw * 8), *w1 = (uint64_t *)k.dev(4, P(rdim, cdim) * nw * 8);
    float *vol = (float *)k.dev(2, P(rdim, cdim) * dispRange * sizeof(float));
    NEED(b0 && b1 && w0 && w1 && vol);
    STEP(rts_memcpy_h2d(b0, censusTrnsRef, P(rdim, cdim) * N, k.s()));
    STEP(rts_memcpy_h2d(b1, censusTrnsTar, P(rdim, cdim) * N, k.s()));
    STEP(rts_census_pack(b0, w0, cdim, rdim, N, k.s()));
    STEP(rts_census_pack(b1, w1, cdim, rdim, N, k.s()));
    STEP(rts_hamming(w0, w1, vol, cdim, rdim, dispRange, nw, k.s()));
    STEP(vol_down(k, vol, cost, cdim, rdim, dispRange));
}

// ---------------------------------------------------------------- rank, LoG, NCC

void calculateRankFilt(unsigned char *img, unsigned char *outImg, const int maskSize, const int rdim, const int cdim) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *a = (uint8_t *)k.dev(0, P(rdim, cdim)), *b = (uint8_t *)k.dev(1, P(rdim, cdim));
    NEED(a && b);
    STEP(rts_memcpy_h2d(a, img, P(rdim, cdim), k.s()));
    STEP(rts_rank(a, b, cdim, rdim, maskSize, k.s()));
    STEP(rts_memcpy_d2h(outImg, b, P(rdim, cdim), k.s()));
    STEP(rts_stream_sync(k.s()));
}

void convolveImage(unsigned char *img, float *outIm, float *mask, const int rdim, const int cdim, const int maskHrad,
                   const int maskWrad) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *a = (uint8_t *)k.dev(0, P(rdim, cdim));
    float *o = (float *)k.dev(1, P(rdim, cdim) * sizeof(float));
    NEED(a && o);
    STEP(rts_memcpy_h2d(a, img, P(rdim, cdim), k.s()));
    STEP(rts_convolve(a, o, mask, cdim, rdim, maskHrad, maskWrad, k.s()));
    STEP(rts_memcpy_d2h(outIm, o, P(rdim, cdim) * sizeof(float), k.s()));
    STEP(rts_stream_sync(k.s()));
}

void calcNCCDispL2R(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                    const int rdim, const int cdim, const int maskHrad, const int maskWrad) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *a = (uint8_t *)k.dev(0, P(rdim, cdim)), *b = (uint8_t *)k.dev(1, P(rdim, cdim));
    float *vol = (float *)k.dev(2, P(rdim, cdim) * dispRange * sizeof(float));
    NEED(a && b && vol);
    STEP(rts_memcpy_h2d(a, ref_image, P(rdim, cdim), k.s()));
    STEP(rts_memcpy_h2d(b, tar_image, P(rdim, cdim), k.s()));
    // The reference leaves the cost untouched where the denominator is 0
    // (seq/NCCDisparity.cpp:70-71): start from the caller's values.
    STEP(vol_up(k, cost, vol, cdim, rdim, dispRange));
    STEP(rts_ncc(a, b, vol, cdim, rdim, dispRange, maskHrad, maskWrad, RTS_NCC_SEQ | RTS_NCC_KEEP_FLAT, k.s()));
    STEP(vol_down(k, vol, cost, cdim, rdim, dispRange));
}

// ---------------------------------------------------------------- combine, box

void combineCost(float *cost1, float *cost2, float lambda1, float lambda2, int rdim, int cdim, int dispRange) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    // Elementwise: the pixel-major layout is used as is.
    const size_t n = P(rdim, cdim) * dispRange;
    float *a = (float *)k.dev(2, n * sizeof(float)), *b = (float *)k.dev(3, n * sizeof(float));
    NEED(a && b);
    STEP(rts_memcpy_h2d(a, cost1, n * sizeof(float), k.s()));
    STEP(rts_memcpy_h2d(b, cost2, n * sizeof(float), k.s()));
    STEP(rts_combine(a, b, lambda1, lambda2, n, k.s()));
    STEP(rts_memcpy_d2h(cost1, a, n * sizeof(float), k.s()));
    STEP(rts_stream_sync(k.s()));
}

void boxAggregate(float *cost, const int rdim, const int cdim, const int dispRange, const int wndwRad) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const size_t vb = P(rdim, cdim) * dispRange * sizeof(float);
    float *a = (float *)k.dev(2, vb), *b = (float *)k.dev(3, vb);
    NEED(a && b);
    STEP(vol_up(k, cost, a, cdim, rdim, dispRange));
    STEP(rts_box(a, b, cdim, rdim, dispRange, wndwRad, k.s()));
    STEP(vol_down(k, b, cost, cdim, rdim, dispRange));
}

// ---------------------------------------------------------------- adaptive weight

void calcProxWeight(float *proxWeight, float gamma_proximity, int maskRad) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const size_t n = (size_t)(2 * maskRad + 1) * (2 * maskRad + 1);
    float *p = (float *)k.dev(0, n * sizeof(float));
    NEED(p);
    STEP(rts_prox_weight(p, gamma_proximity, maskRad, k.s()));
    STEP(rts_memcpy_d2h(proxWeight, p, n * sizeof(float), k.s()));
    STEP(rts_stream_sync(k.s()));
}

// Scalar helpers of the reference (one value, no image data): computed on
// the host in the reference's float order, seq/AdaptiveWeightAggregation.cpp:19-31.
float calcEuclideanDist3P(int l0, int l1, int a0, int a1, int b0, int b1) {
    const float e0 = (float)l0 - (float)l1, e1 = (float)a0 - (float)a1, e2 = (float)b0 - (float)b1;
    const float s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    return sqrtf(s0 + s1 + s2);
}
float calcEuclideanDist3P_float(float l0, float l1, float a0, float a1, float b0, float b1) {
    const float e0 = l0 - l1, e1 = a0 - a1, e2 = b0 - b1;
    const float s0 = e0 * e0, s1 = e1 * e1, s2 = e2 * e2;
    return sqrtf(s0 + s1 + s2);
}

void computeWeights(float *weights, float *proxWeight, unsigned char *LABImage, const float gamma_similarity,
                    const int rdim, const int cdim, const int maskRad) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const size_t A = (size_t)(2 * maskRad + 1) * (2 * maskRad + 1);
    uint8_t *lab = (uint8_t *)k.dev(0, P(rdim, cdim) * 3);
    float *prox = (float *)k.dev(1, A * sizeof(float));
    float *w = (float *)k.dev(2, P(rdim, cdim) * A * sizeof(float));
    NEED(lab && prox && w);
    STEP(rts_memcpy_h2d(lab, LABImage, P(rdim, cdim) * 3, k.s()));
    STEP(rts_memcpy_h2d(prox, proxWeight, A * sizeof(float), k.s()));
    STEP(rts_compute_weights(w, prox, lab, gamma_similarity, cdim, rdim, maskRad, k.s()));
    STEP(rts_memcpy_d2h(weights, w, P(rdim, cdim) * A * sizeof(float), k.s()));
    STEP(rts_stream_sync(k.s()));
}

static void aw_call(const char *fn, float *weightRef, float *weightTar, float *cost, float *out, int rdim, int cdim,
                    int D, int r, int dir) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, fn)) return;
    const size_t A = (size_t)(2 * r + 1) * (2 * r + 1);
    const size_t vb = P(rdim, cdim) * D * sizeof(float), wb = P(rdim, cdim) * A * sizeof(float);
    float *wr = (float *)k.dev(0, wb), *wt = (float *)k.dev(1, wb);
    float *a = (float *)k.dev(2, vb), *b = (float *)k.dev(3, vb);
    if (!wr || !wt || !a || !b) {
        failed(RTS_ERR_NOMEM, fn);
        return;
    }
    if (failed(rts_memcpy_h2d(wr, weightRef, wb, k.s()), fn)) return;
    if (failed(rts_memcpy_h2d(wt, weightTar, wb, k.s()), fn)) return;
    if (failed(vol_up(k, cost, a, cdim, rdim, D), fn)) return;
    if (failed(rts_aw_weights(wr, wt, a, b, cdim, rdim, D, r, dir, k.s()), fn)) return;
    failed(vol_down(k, b, out, cdim, rdim, D), fn);
}

void calcAWCostL2R(float *weightRef, float *weightTar, float *cost, float *newcost, const int rdim, const int cdim,
                   const int dispRange, const int maskRad) {
    aw_call(__func__, weightRef, weightTar, cost, newcost, rdim, cdim, dispRange, maskRad, RTS_L2R);
}
void calcAWCostR2L(float *weightRef, float *weightTar, float *cost, const int rdim, const int cdim, const int dispRange,
                   const int maskRad) {
    aw_call(__func__, weightRef, weightTar, cost, cost, rdim, cdim, dispRange, maskRad, RTS_R2L);
}

// ---------------------------------------------------------------- cross-based

// One pixel's arms (a scalar helper with no volume work): computed on the
// host with the reference's rule, seq/CrossBasedAggregation.cpp:12-100.
static int host_arm(const unsigned char *img, int H, int W, int y, int x, int sy, int sx) {
    const int L = 17, TAO = 20;
    const unsigned char *c = img + 3 * ((size_t)y * W + x);
    int n = 0, delta = 0;
    while (delta < TAO && n < L) {
        const int yy = y + sy * (n + 1), xx = x + sx * (n + 1);
        if (yy < 0 || yy >= H || xx < 0 || xx >= W) break;
        n++;
        const unsigned char *t = img + 3 * ((size_t)yy * W + xx);
        delta = 0;
        for (int ch = 0; ch < 3; ch++) {
            const int v = abs((int)c[ch] - (int)t[ch]);
            if (v > delta) delta = v;
        }
    }
    return n;
}
void calc_x_range(unsigned char *image, int y, int x, int rdim, int cdim, int *leftOffset, int *rightOffset) {
    *leftOffset = host_arm(image, rdim, cdim, y, x, 0, -1);
    *rightOffset = host_arm(image, rdim, cdim, y, x, 0, 1);
}
void calc_y_range(unsigned char *image, int y, int x, int rdim, int cdim, int *topOffset, int *bottomOffset) {
    *topOffset = host_arm(image, rdim, cdim, y, x, -1, 0);
    *bottomOffset = host_arm(image, rdim, cdim, y, x, 1, 0);
}

void findCross(unsigned char *image, int *HorzOffset, int *VertOffset, int rdim, int cdim) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *im = (uint8_t *)k.dev(0, P(rdim, cdim) * 3);
    uint32_t *arms = (uint32_t *)k.dev(1, P(rdim, cdim) * 4);
    int *h = (int *)k.dev(2, P(rdim, cdim) * 8), *v = (int *)k.dev(3, P(rdim, cdim) * 8);
    NEED(im && arms && h && v);
    STEP(rts_memcpy_h2d(im, image, P(rdim, cdim) * 3, k.s()));
    STEP(rts_find_cross(im, arms, cdim, rdim, 17, 20, k.s()));
    STEP(rts_arms_unpack(arms, h, v, cdim, rdim, k.s()));
    STEP(rts_memcpy_d2h(HorzOffset, h, P(rdim, cdim) * 8, k.s()));
    STEP(rts_memcpy_d2h(VertOffset, v, P(rdim, cdim) * 8, k.s()));
    STEP(rts_stream_sync(k.s()));
}

void aggregate_cost_horizontal(float *cost_image, int rdim, int cdim, int dispRange) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    float *a = (float *)k.dev(2, P(rdim, cdim) * dispRange * sizeof(float));
    NEED(a);
    STEP(vol_up(k, cost_image, a, cdim, rdim, dispRange));
    STEP(rts_hprefix(a, cdim, rdim, dispRange, k.s()));
    STEP(vol_down(k, a, cost_image, cdim, rdim, dispRange));
}

void cross_stereo_aggregation(float *cost_image, float *cost_new_image, int *HorzOffset_ref, int *HorzOffset_tar,
                              int *VertOffset_ref, int *VertOffset_tar, int rdim, int cdim, int dispRange) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    const size_t vb = P(rdim, cdim) * dispRange * sizeof(float);
    float *a = (float *)k.dev(2, vb), *b = (float *)k.dev(3, vb);
    uint32_t *ar = (uint32_t *)k.dev(4, P(rdim, cdim) * 4), *at = (uint32_t *)k.dev(5, P(rdim, cdim) * 4);
    NEED(a && b && ar && at);
    STEP(arms_up(k, HorzOffset_ref, VertOffset_ref, ar, cdim, rdim, 6));
    STEP(rts_stream_sync(k.s()));
    STEP(arms_up(k, HorzOffset_tar, VertOffset_tar, at, cdim, rdim, 6));
    STEP(vol_up(k, cost_image, a, cdim, rdim, dispRange));
    int longest = 0;   // arms of at most 17 select the sweep without long-arm fallbacks
    for (const int *m : {HorzOffset_ref, HorzOffset_tar, VertOffset_ref, VertOffset_tar})
        for (size_t i = 0; i < 2 * P(rdim, cdim); i++) longest = std::max(longest, m[i]);
    STEP(rts_cross_v(a, b, ar, at, cdim, rdim, dispRange, RTS_CROSS_EXACT | (longest <= 17 ? RTS_CROSS_ARMS_LE17 : 0),
                     k.s()));
    STEP(vol_down(k, b, cost_new_image, cdim, rdim, dispRange));
}

void disp_refinement(unsigned char *dispIm, int *HorzOffset, int *VertOffset, int rdim, int cdim, int dispRange) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *d = (uint8_t *)k.dev(0, P(rdim, cdim));
    uint32_t *arms = (uint32_t *)k.dev(1, P(rdim, cdim) * 4);
    NEED(d && arms);
    STEP(arms_up(k, HorzOffset, VertOffset, arms, cdim, rdim, 6));
    STEP(rts_memcpy_h2d(d, dispIm, P(rdim, cdim), k.s()));
    // In place in raster order, as the reference (seq/CrossBasedAggregation.cpp:171-212).
    STEP(rts_disp_refine(d, nullptr, arms, cdim, rdim, dispRange, 1, k.s()));
    STEP(rts_memcpy_d2h(dispIm, d, P(rdim, cdim), k.s()));
    STEP(rts_stream_sync(k.s()));
}

// ---------------------------------------------------------------- WTA, cross-check

static void wta_call(const char *fn, float *cost, unsigned char *winners, int rdim, int cdim, int D, int flags) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, fn)) return;
    float *a = (float *)k.dev(2, P(rdim, cdim) * D * sizeof(float));
    uint8_t *dl = (uint8_t *)k.dev(0, P(rdim, cdim));
    if (!a || !dl) {
        failed(RTS_ERR_NOMEM, fn);
        return;
    }
    if (failed(vol_up(k, cost, a, cdim, rdim, D), fn)) return;
    if (failed(rts_wta(a, dl, nullptr, nullptr, cdim, rdim, D, flags, 0, k.s()), fn)) return;
    if (failed(rts_memcpy_d2h(winners, dl, P(rdim, cdim), k.s()), fn)) return;
    failed(rts_stream_sync(k.s()), fn);
}

// The reference's WTA keeps the previous pixel's disparity when d=0 is not
// beaten (seq/WTA.cpp:12-15); the literal mode reproduces it.
void WTA(float *cost, unsigned char *winners, const int rdim, const int cdim, const int dispRange) {
    wta_call(__func__, cost, winners, rdim, cdim, dispRange, RTS_WTA_LITERAL);
}
void WTA_NCC(float *cost, unsigned char *winners, const int rdim, const int cdim, const int dispRange) {
    wta_call(__func__, cost, winners, rdim, cdim, dispRange, RTS_WTA_LITERAL | RTS_WTA_MAX);
}

void cross_check(unsigned char *ref_disparity, unsigned char *tar_disparity, unsigned char *cross_check_image,
                 const int rdim, const int cdim) {
    Call k;
    if (failed(k.ok ? RTS_OK : RTS_ERR_HIP, __func__)) return;
    uint8_t *a = (uint8_t *)k.dev(0, P(rdim, cdim)), *b = (uint8_t *)k.dev(1, P(rdim, cdim)),
            *m = (uint8_t *)k.dev(2, P(rdim, cdim));
    NEED(a && b && m);
    STEP(rts_memcpy_h2d(a, ref_disparity, P(rdim, cdim), k.s()));
    STEP(rts_memcpy_h2d(b, tar_disparity, P(rdim, cdim), k.s()));
    STEP(rts_cross_check(a, b, m, cdim, rdim, 0, k.s()));
    STEP(rts_memcpy_d2h(cross_check_image, m, P(rdim, cdim), k.s()));
    STEP(rts_stream_sync(k.s()));
}
