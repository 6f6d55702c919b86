// rig.h -- device-side description of one 5-view rig and its workspace.
//
// A rig is W x H with max disparity D; every stage works on the replicate-
// padded frame W' = W + 2D, H' = H + 2D (SisterMultiviewDisparities.hpp:35-41).
// Frame F in {0: right, 1: top (90), 2: left (180), 3: bottom (270)} pairs the
// center view with one neighbour.  Frame images are R_F x C_F:
//   F0, F2: H' x W';  F1, F3: W' x H'.
// Frame coordinates (r, c) map to padded original (y, x) as
//   F0: (r, c)            F1: (H'-1-c, W'-1-r)
//   F2: (r, W'-1-c)       F3: (c, W'-1-r)
// (hpp:54-70: flip(.,1), flip(transpose,-1), flip(transpose,0)).
#pragma once
#include <stdint.h>

namespace sister {

constexpr int kMaxDisp = 512;     // packed argmin keys carry d in 9 bits
constexpr int kMaxSide = 16384;   // padded frame side (the serial median holds a row per thread set)
constexpr int kP1 = 7;            // sgm(.., 7, 17, 8) at hpp:280
constexpr int kP2 = 100;          // adaptP2 with alpha = 0, gamma = 100 (sgm.cpp:15-22,34-35)
constexpr int kLrThreshold = 5;   // doLRCheck(.., 5) at hpp:200
constexpr unsigned kGuard = 0x3FFFu;  // stands in for the reference's 65535 guards;
                                      // larger than any path cost + P2, no wrap in u32

struct RigDesc {
    int W, H, D;        // user rig
    int Wp, Hp;         // padded
    int stride;         // row stride of the input views (bytes)
    int Dp;             // cell stride of the D-deep volumes: D rounded up to 16
    const uint8_t* view[5];   // center, right, top, left, bottom (gray u8, W x H)
    uint64_t* code[8];        // census, frame coords: [2F] center frame, [2F+1] neighbour frame
    uint16_t* wta[8];         // raw WTA: [2F] left map, [2F+1] right map (frame coords)
    uint16_t* med[8];         // in-place-median results of wta[]
    uint8_t* mask[4];         // confidence mask of frame F, padded original coords H' x W'
    uint8_t* msum;            // mask-weighted cost sum M, H' x W' x Dp (u8: M <= 4*63)
    uint8_t* part[4];         // per chain family (rows, cols, diag, anti): for rows
                              // 1..H'-2 the two directions' (L - M) summed, in [0,200];
                              // the rows family keeps its forward (L - M) here and its
                              // backward one in rowsb.  Ahead of the fused vertical sweep
                              // only part[0] and rowsb exist (part[1..3] null), and the
                              // sweep's u16 partial sums P replace their bytes row by row
    uint8_t* rowsb;           // rows family's backward (L - M), rows 1..H'-2
    uint32_t* vstate;         // fused vertical sweep: segment-end states, normalised
                              // (L - M) bytes and minL': [2][3][W'][Dp + 16] bytes
    uint16_t* edge;           // exact family contributions on rows 0 and H'-1:
                              // [family][2][W'][Dp]
    uint16_t* sum;            // SGM aggregate S, H' x W' x Dp (debug only, may be null)
    uint16_t* out[3];         // API maps W x H, min(255*d, 65535) (hpp:111-118); nullable
    uint16_t* disp;           // integer multiview disparity W x H; nullable
    float* subpix;            // parabola-refined multiview disparity W x H; nullable
};

__host__ __device__ inline int frame_rows(const RigDesc& r, int F) { return (F & 1) ? r.Wp : r.Hp; }
__host__ __device__ inline int frame_cols(const RigDesc& r, int F) { return (F & 1) ? r.Hp : r.Wp; }

// Deep padding bands: with replicate padding every census code of a frame
// row in [3, D-4] or [R-D+3, R-4] (R frame rows) equals that of the band's
// first row, in the centre and the neighbour frame alike (the 9x7 window sees
// one image row), and so do the row's WTA maps.  band_row gives the row whose
// codes and maps row r holds: the band's first row inside a band, r elsewhere.
// The census stores only first rows, the WTA computes only first rows, and
// readers of either go through band_row.
__host__ __device__ inline int band_row(int r, int R, int D)
{
    return (r > 3 && r <= D - 4) ? 3 : (r > R - D + 3 && r <= R - 4) ? R - D + 3 : r;
}

}  // namespace sister
