/*
 * rtstereo.h -- C ABI of librtstereo, the MI355X (gfx950) dense stereo library.
 *
 * Every entry point takes plain device pointers, sizes and a HIP stream
 * (hipStream_t passed as void*, NULL = the null stream), launches work
 * asynchronously on that stream and returns an rts_status.  Nothing here
 * allocates device memory unless its name says so (rts_malloc, the
 * rts_pipeline_* and rts_server_* objects).
 *
 * Device data layouts
 *   gray image     uint8  [H][W]
 *   colour image   uint8  [H][W][3]  interleaved, as the reference's OpenCV
 *                                    imread buffers (seq/Main.cpp:87-90)
 *   cost volume    float  [D][H][W]  disparity-major: plane d is an H x W
 *                                    image, so a wavefront's 64 lanes touch
 *                                    64 consecutive x of one (d, y) row.
 *                                    The reference's host volumes are
 *                                    pixel-major cost[d + D*(y*W+x)]
 *                                    (seq/PixelBasedCost.cpp:21); rts_vol_to_dm
 *                                    / rts_vol_to_pm convert.
 *   census         uint64 [nw][H][W] bit k of word k/64 is entry k of the
 *                                    reference's bool[H][W][N] descriptor
 *                                    (seq/CensusTransform.cpp:24-28)
 *   arms           uint32 [H][W]     byte 0 left, 1 right, 2 top, 3 bottom
 *                                    (the reference's HorzOffset/VertOffset
 *                                    int pairs, seq/CrossBasedAggregation.cpp:102-115)
 *   disparity      uint8  [H][W]
 *
 * Limits: 1 <= W, H; 1 <= D <= 256 (disparities are uint8); D <= W; window
 * radii smaller than the image side they wrap around.
 *
 * Reference interfaces: each launcher names the seq/ function whose
 * arithmetic it reproduces.  The per-stage C++ drop-in with the reference's
 * exact signatures is declared in rtstereo_seq.h and is built on this ABI.
 */
#ifndef RTSTEREO_H
#define RTSTEREO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int rts_status;
enum {
    RTS_OK = 0,
    RTS_ERR_INVALID = 1,     /* bad argument or unsupported size */
    RTS_ERR_HIP = 2,         /* HIP runtime error */
    RTS_ERR_NOMEM = 3,       /* device allocation failed */
    RTS_ERR_BUDGET = 4,      /* request exceeds the admission budget */
    RTS_ERR_CANCELLED = 5,   /* request cancelled before it ran */
    RTS_ERR_NOTREADY = 6     /* result not ready yet (polling) */
};

/* ------------------------------------------------------------ runtime */

const char *rts_last_error(void);                /* thread-local message of the last failure */
const char *rts_version(void);
rts_status rts_device_count(int *n);
rts_status rts_set_device(int device);
rts_status rts_malloc(void **ptr, size_t bytes);
rts_status rts_free(void *ptr);
rts_status rts_host_alloc(void **ptr, size_t bytes);      /* pinned host memory */
rts_status rts_host_free(void *ptr);
rts_status rts_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
rts_status rts_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
rts_status rts_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
rts_status rts_fill_u8(void *dst, uint8_t value, size_t bytes, void *stream);  /* fill kernel, graph-safe */
rts_status rts_stream_create(void **stream);
rts_status rts_stream_destroy(void *stream);
rts_status rts_stream_sync(void *stream);
rts_status rts_device_sync(void);
rts_status rts_event_create(void **ev);
rts_status rts_event_destroy(void *ev);
rts_status rts_event_record(void *ev, void *stream);
rts_status rts_event_sync(void *ev);
rts_status rts_event_elapsed_ms(void *start, void *stop, float *ms);
rts_status rts_mem_info(size_t *free_bytes, size_t *total_bytes);

/* Pixel-major float volume [H*W][D] <-> disparity-major [D][H][W]. */
rts_status rts_vol_to_dm(const float *pm, float *dm, int W, int H, int D, void *stream);
rts_status rts_vol_to_pm(const float *dm, float *pm, int W, int H, int D, void *stream);

/* Plain copy kernel over `bytes` (multiple of 16), used to calibrate the HBM
 * counters and the achievable copy bandwidth. */
rts_status rts_copy_kernel(const void *src, void *dst, size_t bytes, void *stream);

/* Read-only and write-only kernels of the same shape over `bytes` (multiple
 * of 16): the achievable read and write bandwidths, which bound a stage by
 * the bytes it reads and writes.  The read probe may store into `sink`
 * (4096 words).  No reference counterpart (measurement). */
rts_status rts_read_kernel(const void *src, size_t bytes, void *sink, void *stream);
rts_status rts_write_kernel(void *dst, size_t bytes, void *stream);

/* A float volume of n values (a multiple of 8; both 16-byte aligned) as
 * half, each value rounded to nearest: the FP16 cost volume that
 * rts_aw_half_ws reads.  No reference counterpart (the reference's costs
 * are float). */
rts_status rts_volume_to_half(const float *src, void *dst, size_t n, void *stream);

/* ------------------------------------------------------------ matching costs */

/* Direction of the target pixel: L2R pairs ref x with tar x-d, R2L with x+d;
 * both wrap around the row. */
enum { RTS_L2R = 0, RTS_R2L = 1 };

/* Gray AD |ref - tar|.  seq/PixelBasedCost.cpp:13-37 pixelBasedL2R / pixelBasedR2L. */
rts_status rts_ad_gray(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int dir, void *stream);

/* Float AD.  seq/PixelBasedCost.cpp:39-69 pixelBasedCostL2R_Float / R2L_Float. */
rts_status rts_ad_float(const float *ref, const float *tar, float *vol, int W, int H, int D, int dir, void *stream);

/* Colour AD: sum of the three channel |differences|; truncated to `trunc`
 * when use_trunc; times float(1/3) when norm.  seq/PixelBasedCost.cpp:72-158
 * pixelBasedCostL2R_Color (both overloads) / pixelBasedR2L_Color. */
rts_status rts_ad_color(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int dir,
                        int use_trunc, float trunc, int norm, void *stream);
/* The same costs written as half (each float cost rounded to nearest), the
 * input of rts_aw_half_ws. */
rts_status rts_ad_color_half(const uint8_t *ref, const uint8_t *tar, void *vol_half, int W, int H, int D, int dir,
                             int use_trunc, float trunc, int norm, void *stream);

/* Census descriptors packed into uint64 words ([nw][H][W], nw = ceil(N/64),
 * N = (2hr+1)(2wr+1)).  seq/CensusTransform.cpp:11-37 calcCensusTrans. */
rts_status rts_census(const uint8_t *img, uint64_t *words, int W, int H, int hr, int wr, void *stream);
/* bool[H][W][N] <-> packed words (the reference's census layout). */
rts_status rts_census_unpack(const uint64_t *words, uint8_t *bools, int W, int H, int N, void *stream);
rts_status rts_census_pack(const uint8_t *bools, uint64_t *words, int W, int H, int N, void *stream);

/* Hamming cost popcount(ref(x) ^ tar(x-d)).  seq/CensusTransform.cpp:39-62 calcHammingDist. */
rts_status rts_hamming(const uint64_t *cref, const uint64_t *ctar, float *vol, int W, int H, int D, int nwords,
                       void *stream);

/* Rank transform, radius m.  seq/RankFilter.cpp:11-34 calculateRankFilt. */
rts_status rts_rank(const uint8_t *img, uint8_t *out, int W, int H, int m, void *stream);

/* 2-D float correlation with a (2hr+1)x(2wr+1) host mask (copied to the
 * kernel argument block, at most 15x15 taps), |.|, clamped to 255.
 * seq/ConvolveImage.cpp:11-34 convolveImage. */
rts_status rts_convolve(const uint8_t *img, float *out, const float *mask_host, int W, int H, int hr, int wr,
                        void *stream);

/* NCC cost.  seq/NCCDisparity.cpp:11-73 calcNCCDispL2R.
 * RTS_NCC_SEQ: the reference's float window sums in its order (bit-exact).
 * RTS_NCC_INTEGRAL: exact integer window sums (int64, the denominator in
 *   128-bit integers) combined in FP64, written as float.
 * RTS_NCC_KEEP_FLAT: where the denominator is 0 leave vol untouched (the
 *   reference); otherwise write 0. */
enum { RTS_NCC_SEQ = 0, RTS_NCC_INTEGRAL = 1, RTS_NCC_KEEP_FLAT = 0x100 };
rts_status rts_ncc(const uint8_t *ref, const uint8_t *tar, float *vol, int W, int H, int D, int hr, int wr,
                   int flags, void *stream);

/* In place c1 <- (1-exp(-c1/l1)) + (1-exp(-c2/l2)) over n elements.
 * seq/CombineCost.cpp:11-25 combineCost. */
rts_status rts_combine(float *c1, const float *c2, float l1, float l2, size_t n, void *stream);

/* The three-term cost of rts_pipeline_desc.cost3 in one pass (no reference
 * counterpart; built from seq/RankFilter.cpp + PixelBasedCost.cpp:13-24,
 * ConvolveImage.cpp + PixelBasedCost.cpp:39-53, NCCDisparity.cpp:11-73 and
 * the CombineCost.cpp:11-25 form):
 *   vol = (1-exp(-|rank_ref-rank_tar|/l1)) + (1-exp(-|log_ref-log_tar|/l2))
 *       + (1-exp(-(1-ncc)/l3)),   tar at x-d (wrapped),
 * ncc over a (2hr+1)x(2wr+1) window of the gray images from exact integer
 * sums (0 for a flat window).  hr, wr <= 5. */
rts_status rts_cost_rank_log_ncc(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                 const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol, int W,
                                 int H, int D, int hr, int wr, float l1, float l2, float l3, void *stream);
/* The same with the disparities per block (8, 16 or 32) chosen by the caller
 * (the batch autotuner); rts_cost_rank_log_ncc uses 16. */
rts_status rts_cost_rank_log_ncc_dch(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                     const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol,
                                     int W, int H, int D, int hr, int wr, float l1, float l2, float l3, int dch,
                                     void *stream);
/* Bytes of the workspace of rts_cost_rank_log_ncc_ws at W x H: a 16-byte
 * record per pixel and view. */
size_t rts_cost_rank_log_ncc_ws_bytes(int W, int H);
/* rts_cost_rank_log_ncc_dch with a caller-owned workspace (16-byte aligned,
 * at least rts_cost_rank_log_ncc_ws_bytes(W, H)).  For the 7x7 window
 * (hr = wr = 3) a pre-pass stores each view's per-pixel gray window sums,
 * rank and LoG value there, so the cost kernel forms only the window sums
 * with products of both views per disparity; the output is identical to
 * rts_cost_rank_log_ncc_dch's.  Other windows, or ws == NULL, run
 * rts_cost_rank_log_ncc_dch.  The workspace is in use until the stream
 * reaches the end of the call's work. */
rts_status rts_cost_rank_log_ncc_ws(const uint8_t *gray_ref, const uint8_t *gray_tar, const uint8_t *rank_ref,
                                    const uint8_t *rank_tar, const float *log_ref, const float *log_tar, float *vol,
                                    int W, int H, int D, int hr, int wr, float l1, float l2, float l3, int dch,
                                    void *ws, size_t ws_bytes, void *stream);

/* ------------------------------------------------------------ aggregation */

/* Box mean over (2r+1)^2 with wrapped borders, out of place, float sum in
 * the reference's window order.  seq/ConstBlockAggregation.cpp:11-34 boxAggregate
 * (whose in-place update is replaced by a true box filter). */
rts_status rts_box(const float *in, float *out, int W, int H, int D, int r, void *stream);

/* rts_box for integer-valued volumes, by running sums (O(1) adds per output;
 * r <= 16).  When every input is an integer in [0, 2^24 / (2r+1)^2] (AD of
 * bytes, AD truncated at an integer, census Hamming, rank AD) every partial
 * window sum is exact and the output equals rts_box's bit for bit; other
 * inputs give undefined values.  seq/ConstBlockAggregation.cpp:11-34. */
rts_status rts_box_int(const float *in, float *out, int W, int H, int D, int r, void *stream);

/* Proximity weights exp(-sqrt(dy^2+dx^2)/gp) into a device float[(2r+1)^2].
 * seq/AdaptiveWeightAggregation.cpp:11-18 calcProxWeight. */
rts_status rts_prox_weight(float *prox, float gamma_p, int r, void *stream);

/* Per-pixel support weights float[H][W][(2r+1)^2] (the reference's layout).
 * seq/AdaptiveWeightAggregation.cpp:33-71 computeWeights. */
rts_status rts_compute_weights(float *weights, const float *prox, const uint8_t *lab, float gamma_c, int W, int H,
                               int r, void *stream);

/* Adaptive-weight aggregation from precomputed weights (the reference's
 * dataflow), out of place.  dir RTS_L2R: target centre x-d,
 * seq/AdaptiveWeightAggregation.cpp:73-111 calcAWCostL2R; dir RTS_R2L: the same
 * formula with target centre x+d (the intended form of calcAWCostR2L,
 * seq/AdaptiveWeightAggregation.cpp:113-151, which indexes the weights at the
 * neighbour and updates in place). */
rts_status rts_aw_weights(const float *wref, const float *wtar, const float *vol, float *out, int W, int H, int D,
                          int r, int dir, void *stream);

/* Adaptive-weight aggregation from the two Lab images, without the
 * reference's whole-image weight tensors: seq/AdaptiveWeightAggregation.cpp
 * :11-18 calcProxWeight, :33-71 computeWeights and :73-111 calcAWCostL2R in
 * one call.  The weights of a band of rows go to a workspace, then the band
 * is aggregated.  precision: RTS_AW_FP32 float weights, products and sums
 * (within 1e-5 relative of the FP64 result); RTS_AW_FP16 half weights and
 * products with float accumulation.  r <= 16.  rts_aw allocates the
 * workspace on the stream; rts_aw_ws takes one of at least
 * rts_aw_workspace_bytes(W, H, r, precision) bytes. */
enum { RTS_AW_FP32 = 0, RTS_AW_FP16 = 1 };
rts_status rts_aw(const uint8_t *lab_ref, const uint8_t *lab_tar, const float *vol, float *out, int W, int H, int D,
                  int r, float gamma_c, float gamma_p, int precision, void *stream);
size_t rts_aw_workspace_bytes(int W, int H, int r, int precision);
rts_status rts_aw_ws(const uint8_t *lab_ref, const uint8_t *lab_tar, const float *vol, float *out, int W, int H,
                     int D, int r, float gamma_c, float gamma_p, int precision, void *workspace,
                     size_t workspace_bytes, void *stream);

/* RTS_AW_FP16 aggregation (seq/AdaptiveWeightAggregation.cpp:73-111
 * calcAWCostL2R with the weights of :11-71) from an FP16 cost volume
 * (half [D][H][W], each the float cost rounded to nearest, 16-byte
 * aligned): bit for bit the RTS_AW_FP16 result of rts_aw_ws on the float
 * volume, with up to three disparities per lane so that each weight row
 * read from LDS serves several outputs.  Needs W a multiple of 8 and
 * 384*W*H < 2^32 (rts_aw_half_supported); the workspace is
 * rts_aw_workspace_bytes(W, H, r, RTS_AW_FP16). */
int rts_aw_half_supported(int W, int H, int D);
rts_status rts_aw_half_ws(const uint8_t *lab_ref, const uint8_t *lab_tar, const void *vol_half, float *out, int W,
                          int H, int D, int r, float gamma_c, float gamma_p, void *workspace, size_t workspace_bytes,
                          void *stream);

/* Cross arms (left, right, top, bottom), at most L, colour threshold tao.
 * seq/CrossBasedAggregation.cpp:12-115 calc_x_range / calc_y_range / findCross. */
rts_status rts_find_cross(const uint8_t *rgb, uint32_t *arms, int W, int H, int L, int tao, void *stream);
/* Packed arms <-> the reference's int HorzOffset[H][W][2], VertOffset[H][W][2]. */
rts_status rts_arms_unpack(const uint32_t *arms, int *horz, int *vert, int W, int H, void *stream);
rts_status rts_arms_pack(const int *horz, const int *vert, uint32_t *arms, int W, int H, void *stream);

/* Serial float prefix sum along x of every (d, y) row, in place.
 * seq/CrossBasedAggregation.cpp:117-125 aggregate_cost_horizontal. */
rts_status rts_hprefix(float *vol, int W, int H, int D, void *stream);

/* Vertical sweep of horizontal segments over the prefix volume P: segment
 * sums P[x+r] - P[max(x-l-1,0)] added in the reference's row order, then
 * divided by the count -- bit-identical to
 * seq/CrossBasedAggregation.cpp:127-169 cross_stereo_aggregation.  Both mode
 * values run the same kernel (RTS_CROSS_FAST is kept as an alias).  OR-ing
 * RTS_CROSS_ARMS_LE17 into mode promises that no arm exceeds 17 (the
 * reference's L), which selects a sweep without the long-arm fallbacks. */
enum { RTS_CROSS_EXACT = 0, RTS_CROSS_FAST = 1, RTS_CROSS_ARMS_LE17 = 0x100 };
rts_status rts_cross_v(const float *P, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W, int H,
                       int D, int mode, void *stream);

/* rts_cross_v for arms of at most 17 followed by rts_wta (minimum, the
 * lowest d on ties) and the left-right check, in one pass that never stores
 * the aggregated volume: seq/CrossBasedAggregation.cpp:127-169 then
 * seq/WTA.cpp:10-39 (L2R, and R2L from the volume's diagonal) and
 * seq/CrossCheck.cpp:10-23.  dl, dr and mask equal rts_cross_v + rts_wta's
 * bit for bit.  dr and mask may be null (dl only); with dr, ws holds
 * rts_cross_v_wta_ws_bytes(W, H) bytes (8-byte aligned) of the right view's
 * running bests.  It needs no aggregated volume (1.59 GB less device memory
 * at 1080p/192) but takes longer than the two calls (1.76 against 1.28 ms
 * there): for callers short of memory, not of time.  No reference
 * counterpart as one call. */
size_t rts_cross_v_wta_ws_bytes(int W, int H);
rts_status rts_cross_v_wta(const float *P, const uint32_t *arms_ref, const uint32_t *arms_tar, uint8_t *dl,
                           uint8_t *dr, uint8_t *mask, int W, int H, int D, int tol, void *ws, size_t ws_bytes,
                           void *stream);

/* One full cross-based aggregation from a raw cost volume (no prefix volume
 * in HBM): each segment sum adds its costs in column order, each output adds
 * its rows' segments in row order.  Equal to rts_hprefix + rts_cross_v in
 * exact arithmetic (and bit-equal for integer costs); for float costs it
 * avoids the rounding of the long row prefix, so it is closer to the exact
 * value than the reference's dataflow. */
rts_status rts_cross_fused(const float *cost, float *out, const uint32_t *arms_ref, const uint32_t *arms_tar, int W,
                           int H, int D, void *stream);

/* ------------------------------------------------------------ disparity */

/* Winner-takes-all over d of a float volume, optionally with the right-view
 * disparity read off the volume's diagonal, D_R(x) = argbest_d vol(x+d, d),
 * and the left-right check |D_L(x) - D_R(x-D_L(x))| <= tol in the same pass.
 * seq/WTA.cpp:10-39 WTA / WTA_NCC and seq/CrossCheck.cpp:10-23 cross_check.
 * dr and mask may be NULL (mask requires dr).
 *   RTS_WTA_MAX:     argmax (WTA_NCC) instead of argmin
 *   RTS_WTA_LITERAL: a pixel where no d>0 beats d=0 keeps the previous
 *                    pixel's (raster order) disparity, as seq/WTA.cpp:12-15
 *                    does; otherwise it gets 0.  Applies to D_L only. */
enum { RTS_WTA_MAX = 1, RTS_WTA_LITERAL = 2 };
rts_status rts_wta(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags,
                   int tol, void *stream);

/* rts_wta that allocates nothing: the literal carry's per-row scratch (the
 * only device memory rts_wta uses beyond its arguments) is a fixed array
 * compiled into the library, one per device, of 65535 rows; H above that is
 * rejected with RTS_ERR_INVALID.  Results equal rts_wta's.  Calls on one
 * device share that array, so calls that may overlap on one device (other
 * streams, other threads) race: order them.  No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops). */
rts_status rts_wta_noalloc(const float *vol, uint8_t *dl, uint8_t *dr, uint8_t *mask, int W, int H, int D, int flags,
                           int tol, void *stream);

/* Left-right check of two disparity maps (tol 0 = the reference's equality).
 * seq/CrossCheck.cpp:10-23 cross_check. */
rts_status rts_cross_check(const uint8_t *dl, const uint8_t *dr, uint8_t *mask, int W, int H, int tol, void *stream);

/* Histogram vote over each pixel's cross support.  literal=1 updates in
 * place in raster order (later pixels see refined neighbours), as the
 * reference does; literal=0 reads `disp` and writes `out`.
 * seq/CrossBasedAggregation.cpp:171-212 disp_refinement. */
rts_status rts_disp_refine(uint8_t *disp, uint8_t *out, const uint32_t *arms, int W, int H, int D, int literal,
                           void *stream);

/* ------------------------------------------------------------ pipelines */

/* One stereo pipeline in the order of the reference driver
 * (seq/Main.cpp:180-257): first cost [+ second cost + combine] ->
 * aggregation -> WTA [-> right view + left-right check]. */
enum {
    RTS_COST_AD = 1,      /* colour AD, optional truncation (Main.cpp:181-187) */
    RTS_COST_RANK = 2,    /* rank transform then gray AD (Main.cpp:188-193) */
    RTS_COST_CENSUS = 3,  /* census + Hamming (Main.cpp:194-199) */
    RTS_COST_LOG = 4,     /* LoG then float AD (Main.cpp:200-205) */
    RTS_COST_NCC = 5      /* NCC similarity (Main.cpp:287, commented out there) */
};
enum { RTS_AGGR_NONE = 0, RTS_AGGR_BOX = 1, RTS_AGGR_AW = 2, RTS_AGGR_CROSS = 3 };

typedef struct rts_pipeline_desc {
    int width, height, disp;        /* W, H, D */
    int cost1, cost2;               /* RTS_COST_*; cost2 = 0 for a single cost */
    float lambda1, lambda2;         /* combine weights (cost2 != 0) */
    int ad_trunc_enable;            /* AD truncation on/off */
    float ad_trunc;
    int ad_norm;                    /* AD times 1/3 */
    int census_hr, census_wr;       /* census radii (4, 3 in the reference) */
    int rank_m;                     /* rank radius */
    int ncc_hr, ncc_wr;             /* NCC radii */
    int aggr;                       /* RTS_AGGR_* */
    int box_r;
    int cross_L, cross_tao, cross_iters, cross_exact;
    int aw_r; float aw_gamma_c, aw_gamma_p; int aw_precision;
    int wta_literal;                /* reproduce the reference's WTA carry */
    int lr_check;                   /* compute D_R and the mask */
    int lr_tol;
    /* Third cost term (no counterpart in the reference, which combines at
     * most two costs, seq/CombineCost.cpp:11-25).  Only the combination
     * cost1 = RTS_COST_RANK, cost2 = RTS_COST_LOG, cost3 = RTS_COST_NCC is
     * accepted; the cost is then
     *   (1-exp(-rankAD/lambda1)) + (1-exp(-LoGAD/lambda2)) + (1-exp(-(1-NCC)/lambda3))
     * with NCC = 0 where its window is flat, and it is built in one pass. */
    int cost3;
    float lambda3;
} rts_pipeline_desc;

typedef struct rts_pipeline rts_pipeline;

/* Fills `desc` with the reference driver's constants (seq/Main.cpp:102-131)
 * for the given size. */
void rts_pipeline_defaults(rts_pipeline_desc *desc, int W, int H, int D);
/* Device bytes the pipeline keeps resident (volumes and per-pixel buffers). */
size_t rts_pipeline_bytes(const rts_pipeline_desc *desc);
rts_status rts_pipeline_create(const rts_pipeline_desc *desc, rts_pipeline **out);
rts_status rts_pipeline_destroy(rts_pipeline *p);
/* Runs on device images: colour [H][W][3] each; gray [H][W] is derived on
 * the device when left_gray/right_gray are NULL, and so is Lab [H][W][3]
 * (rts_rgb_to_lab) for adaptive-weight aggregation when left_lab/right_lab
 * are NULL.  dl is required, dr and mask when lr_check. */
rts_status rts_pipeline_run(rts_pipeline *p, const uint8_t *left_rgb, const uint8_t *right_rgb,
                            const uint8_t *left_gray, const uint8_t *right_gray, const uint8_t *left_lab,
                            const uint8_t *right_lab, uint8_t *dl, uint8_t *dr, uint8_t *mask, void *stream);
/* Per-stage timing: when enabled, each run records HIP events on its stream
 * between its stages ("prep": gray/census/arms, "cost": matching cost,
 * "aggr": aggregation, "wta": disparity selection); rts_pipeline_stage_ms
 * waits for the last run and returns the stage durations in ms. */
rts_status rts_pipeline_enable_timing(rts_pipeline *p, int on);
rts_status rts_pipeline_stage_ms(rts_pipeline *p, float *ms, const char **names, int max, int *n);
/* Graph replay: when enabled, the second run with the same image, output
 * and stream pointers captures the run's launch sequence into a HIP graph
 * and every later run with those pointers replays it (new contents in the
 * same buffers are read as usual).  Runs with other pointers, on the null
 * stream or with timing enabled launch their kernels one by one.  For
 * census + cross pipelines the captured graph runs the right view's
 * preparation (gray, packed colour, arms, census) beside the left view's,
 * on a second stream that this call creates.  No reference counterpart (the
 * reference runs on the host). */
rts_status rts_pipeline_enable_graphs(rts_pipeline *p, int on);
/* Device pointer of the final aggregated float volume of the last run
 * (disparity-major), for inspection. */
const float *rts_pipeline_volume(rts_pipeline *p);

/* ------------------------------------------------------------ batches */

/* Many pairs of different sizes under one configuration: the rank + LoG +
 * NCC cost (cost3) with box aggregation (box_r <= 8), WTA, right view and
 * check.  Pairs are cut into chunks whose device footprint
 * (rts_batch_pair_bytes each) fits budget_bytes; in a chunk the box
 * aggregation of all pairs is one persistent launch over a work queue, and
 * with use_graphs the chunk's launch sequence is captured into a HIP graph
 * on first use and replayed when the same pairs (sizes and buffers) run
 * again (a non-null stream is then required).  The first run times the
 * candidate tile shapes (cost disparities per block 8/16/32, box strip rows
 * 32, 64-2r rounded down to a multiple of 4, 64) over the pairs of its first chunk and keeps the fastest (a
 * larger one only when 2% faster).  All buffers are
 * device memory; results are complete when the stream reaches the end of
 * the run.  No reference counterpart (its driver runs one pair per
 * process, seq/Main.cpp:47-299). */
typedef struct rts_batch_item {
    int width, height, disp;
    const uint8_t *left_rgb, *right_rgb;   /* [H][W][3] */
    uint8_t *dl, *dr, *mask;                /* [H][W] */
} rts_batch_item;
typedef struct rts_batch_stats {
    int pairs, chunks;
    int graphs_built, graph_replays;        /* cumulative */
    unsigned long long box_items;           /* cumulative box work items */
    size_t arena_bytes;
    int box_rows, box_cols;                 /* tuned box tile */
    int cost_disparities;                   /* tuned disparities per cost block */
    float tune_ms[9];                       /* cost dch 8/16/32; box tiles 64 x (32, tb, 64)
                                               and 128 x (32, tb, 64), tb = 64-2r rounded
                                               down to a multiple of 4 */
} rts_batch_stats;
typedef struct rts_batch rts_batch;
rts_status rts_batch_create(const rts_pipeline_desc *base, size_t budget_bytes, int use_graphs, rts_batch **out);
rts_status rts_batch_destroy(rts_batch *b);
rts_status rts_batch_run(rts_batch *b, const rts_batch_item *items, int n, void *stream);
/* Fixes the rank + LoG + NCC path's tuned parameters instead of timing the
 * candidates on the first chunk: box tile of box_cols (64 or 128) x
 * box_rows (32, 64, or 64 - 2r rounded down to a multiple of 4) and
 * cost_disparities (8, 16 or 32) per cost block; all three 0 returns to
 * tuning on the next run.  The results do not depend on these choices. */
rts_status rts_batch_set_tiles(rts_batch *b, int box_cols, int box_rows, int cost_disparities);
rts_status rts_batch_get_stats(rts_batch *b, rts_batch_stats *st);
size_t rts_batch_pair_bytes(int W, int H, int D);

/* ------------------------------------------------------------ serving */

/* A server runs stereo requests of mixed size and configuration on `slots`
 * execution slots (one HIP stream and one host worker each), so requests
 * overlap on the device.  Admission is continuous: a queued request starts
 * as soon as a slot is free and its device footprint (rts_pipeline_bytes
 * plus input/output buffers) fits the byte budget; idle pooled pipelines are
 * evicted least-recently-used to make room, and a request that does not fit
 * does not block smaller ones queued behind it.  Requests with stream_id >= 0
 * (camera streams) reuse a pipeline that stays resident, with its buffers,
 * until rts_server_close_stream.  No reference counterpart: the reference
 * driver processes one pair per run (seq/Main.cpp:47-299). */
typedef struct rts_request_desc {
    rts_pipeline_desc pipe;
    int stream_id;                         /* camera stream, or -1 */
    const uint8_t *left_rgb, *right_rgb;   /* host [H][W][3]; valid until the request completes */
    const uint8_t *left_lab, *right_lab;   /* host Lab (adaptive weights), or NULL */
    uint8_t *dl, *dr, *mask;               /* host [H][W] outputs (dr, mask may be NULL without lr_check) */
} rts_request_desc;

typedef struct rts_request_stats {
    double submit_ms, start_ms, done_ms;   /* server clock: queued, started, results in host buffers */
    size_t footprint;                      /* device bytes charged to the budget */
    int status;                            /* rts_status of the request */
} rts_request_stats;

typedef struct rts_server_stats {
    size_t budget_bytes, in_use_bytes;
    int queued, running, resident;
    uint64_t completed;
} rts_server_stats;

typedef struct rts_server rts_server;
/* Each slot runs on a HIP stream of its own.  HIP spreads streams
 * round-robin over GPU_MAX_HW_QUEUES hardware queues (4 by default): a
 * serving process should start with GPU_MAX_HW_QUEUES >= slots + 1 (at most
 * 32), or frames of streams that share a queue run in queue order behind
 * each other's kernels (python -m rtstereo_amd.serve sets 2 x slots). */
rts_status rts_server_create(size_t budget_bytes, int slots, rts_server **out);
rts_status rts_server_destroy(rts_server *s);     /* cancels queued requests, waits for running ones */
/* RTS_ERR_BUDGET if the request alone exceeds the budget. */
rts_status rts_server_submit(rts_server *s, const rts_request_desc *req, uint64_t *ticket);
/* RTS_ERR_NOTREADY until done; then the request's status (stats filled). */
rts_status rts_server_poll(rts_server *s, uint64_t ticket, rts_request_stats *st);
rts_status rts_server_wait(rts_server *s, uint64_t ticket, rts_request_stats *st);
/* Cancels a request that has not started (it then reports RTS_ERR_CANCELLED). */
rts_status rts_server_cancel(rts_server *s, uint64_t ticket);
/* Stops or resumes admission (queued requests wait while paused). */
rts_status rts_server_pause(rts_server *s, int paused);
/* Frames of camera streams (stream_id >= 0) are admitted before one-off
 * requests; at most `limit` one-off requests run at once (default: all
 * slots), which keeps slots free for the streams. */
rts_status rts_server_set_oneoff_limit(rts_server *s, int limit);
rts_status rts_server_close_stream(rts_server *s, int stream_id);
rts_status rts_server_counters(rts_server *s, rts_server_stats *st);

/* Gray conversion used by the pipelines (the reference gets gray from
 * OpenCV's imread, seq/Main.cpp:72-73): gray = (29B + 150G + 77R + 128) >> 8
 * on interleaved B,G,R bytes. */
rts_status rts_rgb_to_gray(const uint8_t *rgb, uint8_t *gray, int W, int H, void *stream);
/* 8-bit CIELab (D65) from B,G,R bytes, for the adaptive weights (the
 * reference takes Lab from OpenCV's cvtColor, seq/Main.cpp:96-97): sRGB
 * linearised, XYZ, then L*255/100, a+128, b+128 rounded and clipped, in FP64. */
rts_status rts_rgb_to_lab(const uint8_t *rgb, uint8_t *lab, int W, int H, void *stream);
/* 3x3 per-channel median with replicated borders on B,G,R bytes, out of
 * place: the OpenCL driver's pre-filter before cross arms
 * (opencl/Main.cpp:342-345). */
rts_status rts_median3(const uint8_t *rgb, uint8_t *out, int W, int H, void *stream);

/* ------------------------------------------------------------ image files */

/* PNG input in the layout of the reference driver's OpenCV imread
 * (seq/Main.cpp:72-73 with flag 0, :87-88 with flag 1): channels = 1 gives
 * gray [H][W], channels = 3 gives B,G,R [H][W][3].  Every PNG colour type and
 * bit depth is read: 16-bit samples keep their high byte, 1/2/4-bit gray
 * levels are scaled to 0..255, palettes are expanded, alpha is dropped,
 * Adam7 interlacing is undone; colour files read as gray take
 * (29B + 150G + 77R + 128) >> 8 (rts_rgb_to_gray's rule).  *pixels is host
 * memory owned by the caller, released with rts_png_free.  Chunk CRCs and
 * the image data's length are checked (RTS_ERR_INVALID on a damaged file). */
rts_status rts_png_read(const char *path, int channels, uint8_t **pixels, int *W, int *H);
/* The same from a PNG file already in memory. */
rts_status rts_png_decode(const uint8_t *data, size_t bytes, int channels, uint8_t **pixels, int *W, int *H);
void rts_png_free(uint8_t *pixels);
/* 8-bit PNG output of a gray [H][W] (channels = 1) or B,G,R [H][W][3]
 * (channels = 3) host image: the reference's imwrite of its disparity image
 * (seq/Main.cpp:275-280). */
rts_status rts_png_write(const char *path, const uint8_t *pixels, int W, int H, int channels);
/* Display scaling of a host disparity map before writing, in place:
 * v <- min(round(v * coeff), 255), with coeff <= 0 meaning 255 / max(v)
 * (a map of zeros is left as it is).  seq/Main.cpp:335-360 scalecvMat. */
rts_status rts_disp_scale(uint8_t *disp, int W, int H, float coeff);

#ifdef __cplusplus
}
#endif
#endif /* RTSTEREO_H */
