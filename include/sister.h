/*
 * sister.h -- C ABI of the MI355X-native SiSter 5-view disparity path.
 *
 * Replaces, for the multiview hot path, the reference's helper class
 *   SisterMultiviewDisparities(center, right, top, left, bottom)
 *     (cpp/include/sister/SisterMultiviewDisparities.hpp:22-24)
 *   void compute_disparities(int dispCount, Mat& multiview, Mat& horizontal,
 *                            Mat& vertical)   (hpp:26-119)
 * and the stage kernels it drives (cpp/include/sister/stereoalgo.h:4-17:
 * ad_census, sgm, WTALeft_SSE, WTARight_SSE, median3x3_SSE, doLRCheck).
 *
 * Plain C: pointers and sizes only.  Every call returns a status code
 * (SISTER_OK = 0); sister_last_error() describes the last failure.  No C++
 * exception crosses this boundary.  A context is not thread safe: use one per
 * host thread, or serialise calls.
 *
 * Images are 8-bit gray, W x H with a row stride in bytes.  W and H must be
 * multiples of 4 and D a multiple of 8 with 8 <= D <= 512 (the reference's own
 * domain is D <= 256: its WTARight_SSE buffer, postprocess.cpp:193; larger D
 * runs here and is flagged by sister_outside_reference_domain()), and the
 * padded sides W + 2D and H + 2D at most 16384.
 * Output maps are W x H uint16 holding min(255 * d, 65535) exactly like the
 * reference's crop-and-scale (hpp:111-118).
 */
#ifndef SISTER_H
#define SISTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    SISTER_OK = 0,
    SISTER_EINVAL = 1,      /* bad argument (shape, D, null pointer) */
    SISTER_ENOMEM = 2,      /* device allocation failed or over the HBM budget */
    SISTER_EDEVICE = 3,     /* HIP runtime error */
    SISTER_EBUSY = 4,       /* serving queue full / admission refused */
    SISTER_ENOTFOUND = 5    /* unknown ticket */
};

/* compute modes (bit m selects the reference's doMultiStereo mode m, hpp:77,82,87) */
enum {
    SISTER_MODE_MULTIVIEW = 1,   /* all four neighbours (mode 0) */
    SISTER_MODE_HORIZONTAL = 2,  /* right + left (mode 1) */
    SISTER_MODE_VERTICAL = 4,    /* top + bottom (mode 2) */
    SISTER_MODE_ALL = 7
};

typedef struct sister_ctx sister_ctx;

typedef struct {
    int device;              /* HIP device ordinal */
    int flags;               /* reserved, 0 */
    size_t hbm_budget;       /* device bytes the context may hold (0: 90% of free HBM) */
} sister_config;

/* One rig.  view[] = {center, right, top, left, bottom}, the constructor
   argument order of the reference (hpp:22).  Output pointers may be NULL
   except out_multiview when SISTER_MODE_MULTIVIEW is requested. */
typedef struct {
    int width, height, stride, disp_count;
    const uint8_t* view[5];
    uint16_t* out_multiview;    /* min(255*d, 65535), W x H */
    uint16_t* out_horizontal;   /* mode 1 map */
    uint16_t* out_vertical;     /* mode 2 map */
    uint16_t* out_disp;         /* integer multiview disparity d */
    float* out_subpix;          /* multiview d + parabola offset (no reference counterpart) */
} sister_rig;

/* hpp:22 + allocation of the device state */
int sister_create(const sister_config* cfg, sister_ctx** out);
void sister_destroy(sister_ctx* ctx);
/* Text of the calling thread's last failed call on this context (like errno:
   callers on several threads each read their own); a thread with no failure
   of its own reads the context's last one.  The serving thread's failures are
   reported through the tickets they fail.  The pointer stays valid until the
   calling thread calls sister_last_error again. */
const char* sister_last_error(const sister_ctx* ctx);

/* compute_disparities (hpp:26-119) for one rig with HOST buffers: uploads,
   runs the requested modes, downloads, returns when the outputs are ready. */
int sister_compute_rig(sister_ctx* ctx, const sister_rig* rig, int modes);

/* Batched form, DEVICE buffers (view[] and outputs in HBM).  All n rigs of
   any sizes run in one launch per stage on the context stream; asynchronous:
   call sister_sync() before reading outputs. */
int sister_run_device(sister_ctx* ctx, const sister_rig* rigs, int n, int modes);

/* sister_run_device without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate device memory (code under
   HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
   The rigs' workspaces and launch tables are carved from a fixed device array
   of SISTER_NOALLOC_MAX_BYTES that each GPU holds from the library's load to
   the end of the process (one 640x480 D=64 rig fits; 1920x1080 D=192 does
   not).  Rigs that need more are refused with SISTER_ENOMEM before anything
   is launched.  Every call on one GPU shares that GPU's array: calls that
   overlap on one GPU (from two contexts, or two threads) race.  Calls on one
   context are ordered: each waits for the context's previous work. */
#define SISTER_NOALLOC_MAX_BYTES ((size_t)256 << 20)
int sister_run_device_noalloc(sister_ctx* ctx, const sister_rig* rigs, int n, int modes);

int sister_sync(sister_ctx* ctx);
/* the hipStream_t the context launches on */
void* sister_stream(sister_ctx* ctx);

/* Device bytes of the per-rig workspace (census codes, maps, cost sum, S). */
size_t sister_workspace_bytes(int width, int height, int disp_count);
/* 1 when D > 256: valid here, outside the reference's working domain */
int sister_outside_reference_domain(int disp_count);

/* ---- capture: the batched run as a replayable hipGraph ------------------ */
/* Captures sister_run_device(rigs, n, modes) into a graph owned by the
   context (replacing any previous one).  Buffers must stay valid. */
int sister_graph_capture(sister_ctx* ctx, const sister_rig* rigs, int n, int modes);
int sister_graph_launch(sister_ctx* ctx);

/* ---- autotuning ----------------------------------------------------------- */
/* No counterpart in the reference (its launch shapes are fixed by its SSE
   loops).  Times on the device, as a captured graph of this batch, the launch
   forms that leave results unchanged: first the graph's parts (1 to 4 parts
   balanced by padded cells, or one part per disparity count, so that groups
   of different D run side by side, and the parts started together or each
   after the previous part's first stages), then for each disparity group of those
   parts the SGM vertical families fused or separate, the lanes per chain of
   the separate families, and the tiled or two-launch cost sum.  The fastest
   of each is kept in the context's table, which every later plan and capture
   of a matching shape consults: groups are matched by D and the half-octaves
   of their padded column total and tallest padded rig, batches by rig count,
   distinct D count, half-octave of padded cells and modes.  Shapes already in
   the table are not timed again.  path (nullable): a forms file, read first
   and rewritten with the table.  The batch is left captured in its tuned
   form (sister_graph_launch replays it). */
int sister_autotune(sister_ctx* ctx, const sister_rig* rigs, int n, int modes, const char* path);
/* The context's table of tuned forms as text, one form per line (the forms
   file's format). */
int sister_tune_describe(sister_ctx* ctx, char* buf, size_t len);

/* ---- serving: continuous batching by HBM footprint ---------------------- */
typedef struct {
    uint64_t submitted, completed, rejected;
    double p50_ms, p99_ms, mean_ms;   /* submit -> result ready, over the last 65536 completions */
    double rigs_per_s;                /* completed / (last completion - first submit) */
    size_t hbm_in_use, hbm_peak;
} sister_stats;

/* Serving takes HOST buffers: a submitted rig is copied in, computed when its
   footprint fits the budget, and its outputs are copied back to the rig's
   output pointers; the caller keeps them valid until sister_wait/poll reports
   completion. */
int sister_serve_start(sister_ctx* ctx, int max_inflight);
int sister_submit(sister_ctx* ctx, const sister_rig* rig, int modes, uint64_t* ticket);
/* The same request with every buffer of the rig in device memory of the
   context's GPU (the ROS callback of ros/README.md:47-56 with its frames
   already on the GPU): nothing is staged or copied; the outputs are written to
   the rig's device buffers, which must stay valid until completion.  The
   footprint charged against the budget is the workspace alone. */
int sister_submit_device(sister_ctx* ctx, const sister_rig* rig, int modes, uint64_t* ticket);
/* A ticket stays known until 65536 later requests have finished; after that
   poll and wait report SISTER_ENOTFOUND for it. */
int sister_poll(sister_ctx* ctx, uint64_t ticket, int* done);
int sister_wait(sister_ctx* ctx, uint64_t ticket);
int sister_serve_stop(sister_ctx* ctx);
int sister_stats_get(sister_ctx* ctx, sister_stats* out);

/* ---- instrumentation and test access ----------------------------------- */
/* Per-kernel device time (HIP events around each kernel's launches on the
   stream they run on) accumulated while enabled: names[i] / ms[i] /
   launches[i] (kernel dispatches issued) for i < returned count.  Profiled
   runs issue the SGM families one after another. */
int sister_profile_enable(sister_ctx* ctx, int on);
int sister_profile_read(sister_ctx* ctx, const char** names, double* ms, int* launches, int max);
/* Copy an intermediate of rig `index` of the last run to host memory:
   what = "census" (F*2+img, u64, frame coords), "wta" / "med" (F*2+side, u16),
   "mask" (F, u8, padded H' x W'), "msum" (u8, H' x W' x Dp), "sum" (u16,
   H' x W' x Dp, only with keep_sum); Dp = D rounded up to a multiple of 16. */
int sister_debug_fetch(sister_ctx* ctx, int index, const char* what, int which, void* dst, size_t bytes);
/* Keep the final aggregate S of every cell in the workspace (by default the
   last sweep consumes it for cells inside the crop). */
int sister_debug_keep_sum(sister_ctx* ctx, int on);
/* Form of the SGM vertical families (columns, diagonals, anti-diagonals) in
   plans built afterwards: 0 = one chain sweep per family, 1 = the fused
   vertical sweep wherever the disparity count allows it (D <= 384),
   -1 = automatic (fused for batches wide enough to fill the device).
   Results are identical either way. */
int sister_debug_vsweep(sister_ctx* ctx, int mode);
/* Plans built afterwards keep the fused vertical sweep's partial sums over
   the rows family's bytes, the smaller layout (2.29 against 3.59 GB per
   1920x1080 D=192 rig, 0.7% slower) that plans take by themselves when they
   do not fit otherwise.  Results are identical either way. */
int sister_debug_sweep_in_place(sister_ctx* ctx, int on);

#ifdef __cplusplus
}
#endif
#endif /* SISTER_H */
