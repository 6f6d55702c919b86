// kernels.h -- launch tables shared by kernels.hip and runtime.hip.
#pragma once
#include <hip/hip_runtime.h>
#include "rig.h"

namespace sister {

enum Stage {
    kStageCensus = 0,
    kStageWta,          // all WTA blocks, or (row form) the rows of frames 0/2
    kStageWtaV,         // row form: the rows of frames 1/3
    kStageWtaFill,      // the WTA maps' deep padding bands, copied from their first rows
    kStageMedianPrep,
    kStageMask,
    kNumStages
};

// Launches whose code depends on the disparity count (cost sum, SGM
// families) run per group of rigs with the same D.
enum GroupKernel {
    kGRows = 0, kGCols, kGDiag, kGAnti,   // SGM families
    kGMsumRows, kGMsumCols,               // cost sum: frames 0/180 (row runs), 90/270 (column runs)
    kGMsumTile,                           // cost sum: all frames, per 16x16 tile
    kGVSweep,                             // fused vertical SGM families, per column block
    kNumGroupKernels
};
struct DGroup {
    const RigDesc* rigs;     // device pointer to the group's first rig
    int nrig;
    int D, Dp;
    const int* bs[kNumGroupKernels];   // block starts per kernel, nrig+1 entries each
    int blocks[kNumGroupKernels];
    int CH[4], KL[4];        // SGM: chunks of 16 disparities per lane, lanes per chain, per family
    int vcols;               // fused vertical sweep: owned columns per block (0: separate families)
    int maxHp;               // tallest padded rig of the group
    int msum_split;          // 1: the two-launch cost sum even where the tiled one fits
};
constexpr int kMaxDGroups = 16;

// Everything one launch of a stage needs for a set of rigs (device pointers).
struct LaunchSet {
    const RigDesc* rigs;
    int nrig;
    const int* bs[kNumStages];      // block-start tables, nrig+1 entries each
    int blocks[kNumStages];         // total blocks per stage
    uint64_t* const* scratch;       // per rig: 8 median order-statistic maps
    int maxD;
    int maxCols;                    // max frame width over the rigs (median LDS)
    int maxWp, maxHp;               // max padded width / height over the rigs
    bool wta_rows;                  // WTA by frame rows from shared costs (k_wta_row)
    int ngroups;
    DGroup groups[kMaxDGroups];
};

// individual kernels, for per-kernel event timing
enum KernelId {
    kKCensus = 0, kKWta, kKMedianPrep, kKMedianSerial, kKMask, kKMsum,
    kKRows, kKCols, kKDiag, kKAnti, kKVSweep, kNumKernels
};
extern const char* const kKernelNames[kNumKernels];

// optional hook called around every kernel launch
struct KernelTimer {
    virtual void begin(int kernel, hipStream_t st) = 0;
    virtual void end(int kernel, hipStream_t st) = 0;
    virtual ~KernelTimer() {}
    long dispatched[kNumKernels] = {};   // kernel dispatches issued, per kernel
};

bool wta_rows_ok(int maxCols, int maxD);
void stage_blocks(const RigDesc& g, int maxD, bool wta_rows, int out[kNumStages]);
int sgm_pick_kl(long cols_chains, int fam, int D);
int sgm_pick_ch(int D, int KL);
int sgm_family_blocks(const RigDesc& g, int KL, int fam);
int msum_blocks(const RigDesc& g, int side);   // side 0 row runs, 1 column runs, 2 tiles
// fused vertical sweep: owned columns per block for a group whose rigs have
// `cols` padded columns in all and disparity count D (0: not used)
int vsweep_pick_cols(long cols, int D, int choice);
bool vsweep_in_place(int D);
int vsweep_blocks(const RigDesc& g, int vcols);
// per-device set-up of kernel attributes (dynamic LDS above 64 KB); call with
// the device current.  0 on success.
int kernels_setup(int device);
void launch_census(const LaunchSet& ls, hipStream_t st, KernelTimer* t);
void launch_wta(const LaunchSet& ls, hipStream_t st, KernelTimer* t);
void launch_median(const LaunchSet& ls, hipStream_t st, KernelTimer* t);
void launch_mask(const LaunchSet& ls, int fmask, hipStream_t st, KernelTimer* t);
void launch_msum(const LaunchSet& ls, int fuse, hipStream_t st, KernelTimer* t);
// Streams that let the rows, columns and diagonals families run concurrently
// (they are independent; the anti-diagonal family, last, reads all three).
struct FamStreams {
    hipStream_t side[2];
    hipEvent_t fork, join[2];
};
// keep_sum: also store the final S of cells inside the crop (debug).
// fs == nullptr runs the four families in sequence on st.
void launch_sgm(const LaunchSet& ls, int mode, int keep_sum, hipStream_t st, const FamStreams* fs, KernelTimer* t);

}  // namespace sister
