/*
 * orchestrate.h -- CPU restatement of SisterMultiviewDisparities::compute_disparities
 * and doMultiStereo (cpp/include/sister/SisterMultiviewDisparities.hpp:26-119,152-295).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).  The orchestration is written once and
 * driven by a table of stage kernels, so the same glue runs either the oracle's
 * own scalar kernels (oracle/sister_oracle.c) or the reference's compiled
 * kernels (oracle/_ref, built from census.cpp / sgm.cpp / postprocess.cpp).
 */
#ifndef SISTER_ORACLE_ORCHESTRATE_H
#define SISTER_ORACLE_ORCHESTRATE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    /* ad_census (census.cpp:149-158): census of both frame images + Hamming
       cost volume [rows][cols][D]; dsi arrives zero-filled */
    void (*cost)(const uint8_t* imc, const uint8_t* imn, int rows, int cols, int D, uint16_t* dsi);
    /* censusTransform (census.cpp:38-51), rw=4 rh=3; may be NULL */
    void (*census)(const uint8_t* img, int rows, int cols, uint64_t* code);
    void (*wta_left)(uint16_t* dsi, int rows, int cols, int D, float* out);
    void (*wta_right)(uint16_t* dsi, int rows, int cols, int D, float* out);
    void (*median_inplace)(float* buf, int rows, int cols);
    void (*lrcheck)(float* L, float* R, int rows, int cols, int thr);
    /* sgm (sgm.cpp:457-460); img = padded center view, H x W */
    void (*sgm)(uint16_t* dsi, const uint8_t* img, int H, int W, int D, uint16_t* S);
} sister_cpu_kernels;

/* Optional intermediates.  Frame order everywhere: 0 = right, 1 = top (90),
   2 = left (180), 3 = bottom (270).  Frame images are R_F x C_F with
   (R,C) = (H',W') for frames 0/2 and (W',H') for frames 1/3. */
typedef struct {
    uint64_t* census[8];   /* [2F] center frame, [2F+1] neighbour frame; frame coords */
    float* wta_l[4];       /* raw WTALeft per frame (frame coords) */
    float* wta_r[4];       /* raw WTARight per frame */
    float* med_l[4];       /* after the in-place median */
    float* med_r[4];
    uint8_t* mask[4];      /* confidence mask mapped to padded original coords H' x W' */
    uint16_t* msum[3];     /* per mode: mask-weighted cost sum H' x W' x D */
    uint16_t* sum[3];      /* per mode: SGM aggregate S, H' x W' x D */
    float* disp_pad[3];    /* per mode: WTALeft on S, H' x W' */
} sister_cpu_debug;

/* compute_disparities: views = gray u8 {center, right, top, left, bottom},
   W x H with row stride `stride`.  modes: bit m set computes mode m
   (0 = multiview, 1 = horizontal, 2 = vertical).  out[m] = W x H uint16
   holding min(255*d, 65535) (hpp:111-118).  subpix (nullable) receives the
   parabola-refined multiview disparity (mode 0 must be requested).
   disp (nullable) receives the integer multiview disparity. */
int sister_cpu_compute(const sister_cpu_kernels* k, const uint8_t* const views[5],
                       int W, int H, int stride, int D, int modes,
                       uint16_t* out[3], float* subpix, uint16_t* disp,
                       sister_cpu_debug* dbg);

#ifdef __cplusplus
}
#endif
#endif
