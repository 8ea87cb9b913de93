/*
 * rtstereo_seq.h -- drop-in replacement for the reference's sequential stage
 * functions (haibo031031/rtstereo, the seq/ headers).  Same names, C++ linkage,
 * same argument order and meaning, same host buffers and layouts; the work
 * runs on the GPU through librtstereo's C ABI (rtstereo.h).  Link
 * librtstereo_seq.so instead of compiling the reference's seq/ stage files.
 *
 * Conventions kept from the reference: rdim = image height, cdim = width,
 * dispRange = number of disparities (<= 256), radii rather than widths, cost
 * volumes pixel-major cost[d + dispRange*(y*cdim + x)], every window wraps
 * around the image borders (except adaptive weights and cross arms, which
 * stop at them).  All calls are synchronous: results are in the caller's
 * buffers when the function returns.
 *
 * Behaviour that differs from the reference on purpose:
 *   - boxAggregate computes a true box mean (the reference updates the
 *     volume while it reads it, seq/ConstBlockAggregation.cpp:27,31);
 *   - calcAWCostR2L uses the L2R formula with the target centre at x+d
 *     (the reference's version reads its weights at the neighbour and
 *     updates in place, seq/AdaptiveWeightAggregation.cpp:140-148).
 * Errors (bad sizes, device failures) print a message and abort, as the
 * reference's GPU versions do; set RTS_SEQ_NOABORT=1 to return instead and
 * read rts_seq_last_error().
 */
#ifndef RTSTEREO_SEQ_H
#define RTSTEREO_SEQ_H

/* seq/PixelBasedCost.h:6-13 */
void pixelBasedL2R(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange, const int rdim,
                   const int cdim);
void pixelBasedR2L(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange, const int rdim,
                   const int cdim);
void pixelBasedCostL2R_Float(float *ref_image, float *tar_image, float *cost, const int dispRange, const int rdim,
                             const int cdim);
void pixelBasedCostR2L_Float(float *ref_image, float *tar_image, float *cost, const int dispRange, const int rdim,
                             const int cdim);
void pixelBasedCostL2R_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, int normOrNot);
void pixelBasedCostL2R_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, float trunc, int normOrNot);
/* Declared as pixelBasedCostR2L_Color in seq/PixelBasedCost.h:11,13 but
 * defined as pixelBasedR2L_Color in seq/PixelBasedCost.cpp:116,138: both
 * names are exported. */
void pixelBasedCostR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, int normOrNot);
void pixelBasedCostR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                             const int rdim, const int cdim, float trunc, int normOrNot);
void pixelBasedR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                         const int rdim, const int cdim, int normOrNot);
void pixelBasedR2L_Color(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                         const int rdim, const int cdim, float trunc, int normOrNot);

/* seq/CensusTransform.h:1-2 */
void calcCensusTrans(unsigned char *img, bool *censusTrans, const int maskHrad, const int maskWrad, const int rdim,
                     const int cdim);
void calcHammingDist(bool *censusTrnsRef, bool *censusTrnsTar, float *cost, const int maskHrad, const int maskWrad,
                     int rdim, int cdim, int dispRange);

/* seq/RankFilter.h:1 (declaration restated from seq/RankFilter.cpp:11) */
void calculateRankFilt(unsigned char *img, unsigned char *outImg, const int maskSize, const int rdim, const int cdim);

/* seq/ConvolveImage.h:1 (declaration restated from seq/ConvolveImage.cpp:11) */
void convolveImage(unsigned char *img, float *outIm, float *mask, const int rdim, const int cdim, const int maskHrad,
                   const int maskWrad);

/* seq/NCCDisparity.h:3 */
void calcNCCDispL2R(unsigned char *ref_image, unsigned char *tar_image, float *cost, const int dispRange,
                    const int rdim, const int cdim, const int maskHrad, const int maskWrad);

/* seq/CombineCost.h:3 */
void combineCost(float *cost1, float *cost2, float lambda1, float lambda2, int rdim, int cdim, int dispRange);

/* seq/ConstBlockAggregation.h:2 */
void boxAggregate(float *cost, const int rdim, const int cdim, const int dispRange, const int wndwRad);

/* seq/AdaptiveWeightAggreagtion.h:4-11 */
void calcProxWeight(float *proxWeight, float gamma_proximity, int maskRad);
float calcEuclideanDist3P(int l0, int l1, int a0, int a1, int b0, int b1);
float calcEuclideanDist3P_float(float l0, float l1, float a0, float a1, float b0, float b1);
void computeWeights(float *weights, float *proxWeight, unsigned char *LABImage, const float gamma_similarity,
                    const int rdim, const int cdim, const int maskRad);
void calcAWCostL2R(float *weightRef, float *weightTar, float *cost, float *newcost, const int rdim, const int cdim,
                   const int dispRange, const int maskRad);
void calcAWCostR2L(float *weightRef, float *weightTar, float *cost, const int rdim, const int cdim, const int dispRange,
                   const int maskRad);

/* seq/CrossBasedAggregation.h:13-20 (medianFilt at :12 is declared but never
 * defined by the reference and is not provided) */
void calc_x_range(unsigned char *image, int y, int x, int rdim, int cdim, int *leftOffset, int *rightOffset);
void calc_y_range(unsigned char *image, int y, int x, int rdim, int cdim, int *topOffset, int *bottomOffset);
void findCross(unsigned char *image, int *HorzOffset, int *VertOffset, int rdim, int cdim);
void aggregate_cost_horizontal(float *cost_image, int rdim, int cdim, int dispRange);
void cross_stereo_aggregation(float *cost_image, float *cost_new_image, int *HorzOffset_ref, int *HorzOffset_tar,
                              int *VertOffset_ref, int *VertOffset_tar, int rdim, int cdim, int dispRange);
void disp_refinement(unsigned char *dispIm, int *HorzOffset, int *VertOffset, int rdim, int cdim, int dispRange);

/* seq/WTA.h:1-2 */
void WTA(float *cost, unsigned char *winners, const int rdim, const int cdim, const int dispRange);
void WTA_NCC(float *cost, unsigned char *winners, const int rdim, const int cdim, const int dispRange);

/* seq/CrossCheck.h:1 (declaration restated from seq/CrossCheck.cpp:10) */
void cross_check(unsigned char *ref_disparity, unsigned char *tar_disparity, unsigned char *cross_check_image,
                 const int rdim, const int cdim);

/* Not in the reference: message of the last failed call on this thread
 * (empty when the last call succeeded). */
extern "C" const char *rts_seq_last_error(void);

#endif /* RTSTEREO_SEQ_H */
