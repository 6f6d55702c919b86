/*
 * ref_glue.cpp -- binds the reference's own CPU kernels (census.cpp, sgm.cpp,
 * postprocess.cpp, compiled unmodified from the reference checkout by
 * oracle/Makefile into oracle/_ref/libsister_ref.so) to the shared
 * orchestration in orchestrate.c.
 *
 * TEST INFRASTRUCTURE ONLY: used to pin the oracle (golden vectors) and as the
 * "reference" CPU baseline timed by bench.py.  Nothing here is product code.
 *
 * The entry points called are exactly the ones the helper class calls
 * (SisterMultiviewDisparities.hpp:181-184,196-200,280,283), with the same
 * arguments: ad_census(..., numThreads=4), WTALeft_SSE/WTARight_SSE with
 * uniqueness 1, median3x3_SSE in place, doLRCheck(.., 5), sgm(.., 7, 17, 8).
 */
#include "stereoalgo.h"
#include "orchestrate.h"

/* defined with external linkage in census.cpp:38, not declared in stereoalgo.h */
void censusTransform(uint8* source, uint64* dest, uint32 width, uint32 height, uint8 rw, uint8 rh);

static void r_cost(const uint8_t* imc, const uint8_t* imn, int rows, int cols, int D, uint16_t* dsi)
{
    ad_census((uint8*)imc, (uint8*)imn, rows, cols, D, dsi, 4);
}
static void r_census(const uint8_t* img, int rows, int cols, uint64_t* code)
{
    censusTransform((uint8*)img, (uint64*)code, (uint32)cols, (uint32)rows, 4, 3);
}
static void r_wta_left(uint16_t* dsi, int rows, int cols, int D, float* out)
{
    WTALeft_SSE(out, dsi, cols, rows, D - 1, 1);
}
static void r_wta_right(uint16_t* dsi, int rows, int cols, int D, float* out)
{
    WTARight_SSE(out, dsi, cols, rows, D - 1, 1);
}
static void r_median(float* buf, int rows, int cols)
{
    median3x3_SSE(buf, buf, (uint32)cols, (uint32)rows);
}
static void r_lrcheck(float* L, float* R, int rows, int cols, int thr)
{
    doLRCheck(L, R, cols, rows, thr);
}
static void r_sgm(uint16_t* dsi, const uint8_t* img, int H, int W, int D, uint16_t* S)
{
    sgm((uint8*)img, H, W, D, dsi, S, 7, 17, 8);   /* hpp:280 */
}

static const sister_cpu_kernels ref_kernels = {
    r_cost, r_census, r_wta_left, r_wta_right, r_median, r_lrcheck, r_sgm,
};

extern "C" __attribute__((visibility("default")))
int ref_compute(const uint8_t* const views[5], int W, int H, int stride, int D, int modes,
                uint16_t* out[3], float* subpix, uint16_t* disp, sister_cpu_debug* dbg)
{
    if (D > 256) return -2;  /* WTARight_SSE's store[256+32] (postprocess.cpp:193) */
    return sister_cpu_compute(&ref_kernels, views, W, H, stride, D, modes, out, subpix, disp, dbg);
}

extern "C" __attribute__((visibility("default")))
void ref_census(const uint8_t* img, int rows, int cols, uint64_t* code) { r_census(img, rows, cols, code); }

/* The stage kernels one by one, for glue written elsewhere (the numpy
   restatement of the helper class in tests/glue_np.py drives these). */
extern "C" __attribute__((visibility("default")))
void ref_cost(const uint8_t* imc, const uint8_t* imn, int rows, int cols, int D, uint16_t* dsi)
{
    r_cost(imc, imn, rows, cols, D, dsi);
}
extern "C" __attribute__((visibility("default")))
void ref_wta_left(uint16_t* dsi, int rows, int cols, int D, float* out) { r_wta_left(dsi, rows, cols, D, out); }
extern "C" __attribute__((visibility("default")))
void ref_wta_right(uint16_t* dsi, int rows, int cols, int D, float* out) { r_wta_right(dsi, rows, cols, D, out); }
extern "C" __attribute__((visibility("default")))
void ref_median(float* buf, int rows, int cols) { r_median(buf, rows, cols); }
extern "C" __attribute__((visibility("default")))
void ref_lrcheck(float* L, float* R, int rows, int cols, int thr) { r_lrcheck(L, R, rows, cols, thr); }
extern "C" __attribute__((visibility("default")))
void ref_sgm(uint16_t* dsi, const uint8_t* img, int H, int W, int D, uint16_t* S) { r_sgm(dsi, img, H, W, D, S); }
