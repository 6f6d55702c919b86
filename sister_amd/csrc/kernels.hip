// kernels.hip -- gfx950 kernels of the 5-view disparity path.
//
// Stage map (reference file:line -> kernel):
//   censusTransform        census.cpp:38-51             k_census
//   WTALeft/WTARight on    postprocess.cpp:74-315       k_wta   (cost recomputed from
//     the raw frame DSIs     (hpp:196-197,...)                   census, never stored)
//   median3x3_SSE in place postprocess.cpp:15-71        k_median_prep + k_median_serial
//   doLRCheck + mask loops postprocess.cpp:318-341,     k_mask
//                          hpp:201-251
//   realign + weighted sum hpp:255-277                  k_msum
//   accumulateCostsSSE     sgm.cpp:26-455               k_fam<.., 0..3> (rows, columns,
//                                                       diagonals, anti-diagonals)
//   WTALeft on S, convert  hpp:283,111-118              fused into k_fam<.., 3>
//
// Every kernel takes the rig table plus a block-start table so one launch can
// cover many rigs of different sizes: block b belongs to rig i when
// bs[i] <= b < bs[i+1].
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include "rig.h"
#include "kernels.h"

namespace sister {

// Device buffers reached through the rig table are generic pointers; casting
// them to the global address space lets the compiler use global_load/store
// with a scalar base (flat accesses need 64-bit per-lane addresses and also
// hold the LDS counter).
template <typename T> using gptr = __attribute__((address_space(1))) T*;
template <typename T> __device__ __forceinline__ gptr<T> G(T* p) { return (gptr<T>)p; }
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 LD16(const void* p)
{
    const u32x4 v = *(gptr<const u32x4>)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void ST16(void* p, uint4 v) { *(gptr<u32x4>)p = u32x4{v.x, v.y, v.z, v.w}; }

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
// Block index b of an n-block grid taken in XCD-contiguous order: with
// round-robin dispatch block b runs on XCD b % 8, so the XCD gets the
// consecutive logical blocks [start(b % 8), start(b % 8 + 1)) (the first
// n % 8 XCDs one more than the others).  A bijection of [0, n).
__device__ __forceinline__ int xcd_order(int b, int n)
{
    const int x = b & 7, k = b >> 3, q = n >> 3, r = n & 7;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

__device__ __forceinline__ int find_rig(const int* __restrict__ bs, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (bs[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

// Popcounts accumulate in place: v_bcnt_u32_b32 adds its second operand, one
// instruction per 32-bit half (written out, since the compiler reassociates
// pop + pop + acc into two plain popcounts and a three-way add).
__device__ __forceinline__ unsigned bcnt_add(uint32_t x, unsigned acc)
{
    unsigned r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

// padded original (y, x) of frame pixel (r, c)
__device__ __forceinline__ void frame_to_orig(const RigDesc& g, int F, int r, int c, int& y, int& x)
{
    switch (F) {
    case 0: y = r; x = c; break;
    case 1: y = g.Hp - 1 - c; x = g.Wp - 1 - r; break;
    case 2: y = r; x = g.Wp - 1 - c; break;
    default: y = c; x = g.Wp - 1 - r; break;
    }
}

// frame pixel (r, c) of padded original (y, x)
__device__ __forceinline__ void orig_to_frame(const RigDesc& g, int F, int y, int x, int& r, int& c)
{
    switch (F) {
    case 0: r = y; c = x; break;
    case 1: r = g.Wp - 1 - x; c = g.Hp - 1 - y; break;
    case 2: r = y; c = g.Wp - 1 - x; break;
    default: r = g.Wp - 1 - x; c = y; break;
    }
}

// pixel of the replicate-padded view (copyMakeBorder REPLICATE, hpp:35-39)
__device__ __forceinline__ int padded_px(const RigDesc& g, const uint8_t* v, int y, int x)
{
    y = min(max(y - g.D, 0), g.H - 1);
    x = min(max(x - g.D, 0), g.W - 1);
    return v[(size_t)y * g.stride + x];
}

__device__ __forceinline__ int frame_px(const RigDesc& g, const uint8_t* v, int F, int r, int c)
{
    int y, x;
    frame_to_orig(g, F, r, c, y, x);
    return padded_px(g, v, y, x);
}

// cross-lane moves (DPP).  shr1: lane i gets lane i-1, lane 0 gets `fill`.
// shl1: lane i gets lane i+1, lane 63 gets `fill`.
__device__ __forceinline__ unsigned shr1(unsigned v, unsigned fill)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x138, 0xF, 0xF, false);
}
__device__ __forceinline__ unsigned shl1(unsigned v, unsigned fill)
{
    return (unsigned)__builtin_amdgcn_update_dpp((int)fill, (int)v, 0x130, 0xF, 0xF, false);
}

// minimum over the 64 lanes, returned wave-uniform
__device__ __forceinline__ unsigned wave_min(unsigned v)
{
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x111, 0xF, 0xF, false));  // row_shr:1
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x112, 0xF, 0xF, false));  // row_shr:2
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x114, 0xF, 0xF, false));  // row_shr:4
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x118, 0xF, 0xF, false));  // row_shr:8
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x142, 0xA, 0xF, false));  // row_bcast:15
    v = min(v, (unsigned)__builtin_amdgcn_update_dpp(-1, (int)v, 0x143, 0xC, 0xF, false));  // row_bcast:31
    return (unsigned)__builtin_amdgcn_readlane((int)v, 63);
}

// ---------------------------------------------------------------------------
// K1 census (census.cpp:38-51), one 32x32 tile of frame pixels per block.
// Test bit for offset (y,x), y=-3..3 outer, x=-4..4 inner, is
// I(r+y, c+x) > I(r-y, c-x), shifted in MSB-first; bit 63 carries bit 0 of
// the previous domain pixel in frame raster order (the accumulator is never
// reset, census.cpp:42), i.e. the (3,4) test of that pixel.
// The tile (frame rows r0-3 .., columns c0-8 ..) is staged in LDS as bytes,
// rows 40 dwords apart (40 = 8 mod 32 banks: the 32 lanes of a half-wave,
// 4 rows x 8 runs, read 32 different banks).  A thread takes a run of 4
// frame pixels of one row and reads each of its 7 window rows once as 3 or 4
// aligned dwords; one alignbyte gathers a test's 4 left-hand bytes (one per
// pixel), another the 4 right-hand bytes, and each pixel's two bits come from
// two byte-select subtractions.
// ---------------------------------------------------------------------------
constexpr int kCT = 32;   // tile side (frame pixels)

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) void k_census(const RigDesc* __restrict__ rigs, int nrig,
                                                const int* __restrict__ bs)
{
    constexpr int TR = kCT + 6;       // frame rows r0-3 .. r0+kCT+2
    constexpr int TB = kCT + 12;      // bytes used per row: columns c0-8 .. c0+kCT+3
    constexpr int TW = 40;            // dwords per tile row
    __shared__ uint32_t tile[TR * TW];
    uint8_t* const tb = (uint8_t*)tile;
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lb = blockIdx.x - bs[ri];
    const int T = ((g.Wp + kCT - 1) / kCT) * ((g.Hp + kCT - 1) / kCT);
    const int unit = lb / T, t = lb % T;
    const int F = unit >> 1;
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const int tilesC = (C + kCT - 1) / kCT;
    const int r0 = (t / tilesC) * kCT, c0 = (t % tilesC) * kCT;
    const uint8_t* v = g.view[(unit & 1) ? 1 + F : 0];
    // rows of the deep padding bands other than their first rows are never
    // stored (readers take the first row, band_row): a tile of such rows only
    // has nothing to do
    if (r0 + kCT - 1 < R && band_row(r0, R, g.D) != r0 && band_row(r0 + kCT - 1, R, g.D) != r0 + kCT - 1 &&
        band_row(r0, R, g.D) == band_row(r0 + kCT - 1, R, g.D))
        return;

    // frames 90/270 are transposes: step along frame rows there, so that
    // consecutive lanes read along an image row
    const bool odd = F & 1;
    for (int i = threadIdx.x; i < TR * TB; i += 256) {
        const int rr = odd ? i % TR : i / TB, cc = odd ? i / TR : i % TB;
        const int r = r0 - 3 + rr, c = c0 - 8 + cc;
        tb[rr * TW * 4 + cc] = (r >= 0 && r < R && c >= 0 && c < C) ? (uint8_t)frame_px(g, v, F, r, c) : 0;
    }
    __syncthreads();
    // codes of the run of pixels (r0 + rl, c0 + 4 run .. + 3), from the tile
    auto run_codes = [&](int rl, int run, uint64_t (&code)[4]) __attribute__((always_inline)) {
        const int r = r0 + rl, cb = c0 + 4 * run;
        // w[i][k]: dword run + k of tile row rl + i (frame row r - 3 + i); byte b of
        // the run's window is frame column cb - 8 + b
        uint32_t w[7][4];
#pragma unroll
        for (int i = 0; i < 7; i++)
#pragma unroll
            for (int k = 0; k < 4; k++) w[i][k] = tile[(rl + i) * TW + run + k];
        // bytes s .. s+3 of row i: byte p is pixel p's entry at window byte s + p
        auto quad = [&](int i, int sb) __attribute__((always_inline)) {
            return (sb & 3) == 0 ? w[i][sb >> 2] : __builtin_amdgcn_alignbyte(w[i][(sb >> 2) + 1], w[i][sb >> 2], sb & 3);
        };
        // test u (offset (u/9 - 3, u%9 - 4)) and its mirror 62 - u compare the same
        // two pixels in opposite order: bit 62 - u = (a > b) is the sign of b - a,
        // bit u = (b > a) the sign of a - b.  hi collects bits 62..32 MSB-first;
        // lf collects bits 0..30 in the same order and is bit-reversed at the end.
        uint32_t hi[4] = {0, 0, 0, 0}, lf[4] = {0, 0, 0, 0};
#pragma unroll
        for (int u = 0; u < 31; u++) {
            const int y = u / 9 - 3, x = u % 9 - 4;
            const uint32_t qa = quad(3 + y, 8 + x), qb = quad(3 - y, 8 - x);
#pragma unroll
            for (int p = 0; p < 4; p++) {
                const uint32_t a = (qa >> (8 * p)) & 0xFFu, b = (qb >> (8 * p)) & 0xFFu;
                hi[p] = __builtin_amdgcn_alignbit(hi[p], b - a, 31);
                lf[p] = __builtin_amdgcn_alignbit(lf[p], a - b, 31);
            }
        }
        // carry of pixel c: the (3,4) test of pixel c-1, I(r+3, c+3) > I(r-3, c-5)
        const uint32_t ca = quad(6, 11 + 0), cbq = quad(0, 3);
#pragma unroll
        for (int p = 0; p < 4; p++) {
            const int c = cb + p;
            code[p] = 0;
            if (r >= 3 && r < R - 3 && c >= 4 && c < C - 4) {
                const uint32_t lo = __builtin_bitreverse32(lf[p]) >> 1;
                unsigned carry;
                if (c > 4)
                    carry = ((ca >> (8 * p)) & 0xFFu) > ((cbq >> (8 * p)) & 0xFFu);
                else if (r > 3)
                    carry = frame_px(g, v, F, r + 2, C - 1) > frame_px(g, v, F, r - 4, C - 9);  // (r-1, C-5)
                else
                    carry = 0;
                code[p] = ((uint64_t)(hi[p] | (carry << 31)) << 32) | lo;
            }
        }
    };
    auto store_run = [&](int r, int cb, const uint64_t (&code)[4]) __attribute__((always_inline)) {
        if (band_row(r, R, g.D) != r) return;                // a band's copy row: not stored
        uint64_t* out = g.code[unit] + (size_t)r * C + cb;   // 32-byte aligned: cb % 4 == 0
        ST16(out, make_uint4((uint32_t)code[0], (uint32_t)(code[0] >> 32), (uint32_t)code[1], (uint32_t)(code[1] >> 32)));
        ST16(out + 2, make_uint4((uint32_t)code[2], (uint32_t)(code[2] >> 32), (uint32_t)code[3], (uint32_t)(code[3] >> 32)));
    };
    const int run = threadIdx.x & 7, rl = threadIdx.x >> 3;
    // Tiles in the deep padding.  With replicate padding the window of a pixel
    // at frame rows [3, D-4] holds the image's first row only (rows
    // [R-D+3, R-4]: its last row), so its code does not depend on the row; the
    // carry of column 4 compares two copies of one pixel there (0, as at row
    // 3).  Likewise a pixel at columns [5, D-5] or [C-D+4, C-5] sees one image
    // column and its carry compares two copies of the same column, so within a
    // row its code does not depend on the column.  A tile wholly inside such a
    // band computes its distinct codes in one wave (one run of 4 per tile row
    // band, one pixel per row for a column band, one code for both) and copies.
    const int D = g.D;
    const bool rows_deep = (r0 >= 3 && r0 + kCT - 1 <= D - 4) || (r0 >= R - D + 3 && r0 + kCT - 1 <= R - 4);
    const bool cols_deep = (c0 >= 5 && c0 + kCT - 1 <= D - 5) || (c0 >= C - D + 4 && c0 + kCT - 1 <= C - 5);
    const bool deep = rows_deep || cols_deep;   // uniform over the block
    const int nrep = rows_deep ? (cols_deep ? 1 : kCT / 4) : kCT;
    const bool work = deep ? (int)threadIdx.x < nrep : r0 + rl < R && c0 + 4 * run < C;
    uint64_t code[4] = {0, 0, 0, 0};
    if (work)
        run_codes(deep ? (rows_deep ? 0 : (int)threadIdx.x) : rl,
                  deep ? (rows_deep && !cols_deep ? (int)threadIdx.x : 0) : run, code);
    if (deep) {
        __shared__ uint64_t rep[kCT * 4];
        if (work) {
#pragma unroll
            for (int p = 0; p < 4; p++) rep[threadIdx.x * 4 + p] = code[p];
        }
        __syncthreads();
        // a row band's tile may end past column C, a column band's past row R
        if (r0 + rl >= R || c0 + 4 * run >= C) return;
#pragma unroll
        for (int p = 0; p < 4; p++)
            code[p] = rows_deep ? (cols_deep ? rep[0] : rep[run * 4 + p]) : rep[rl * 4];
    } else if (!work) {
        return;   // C is a multiple of 4, so a run is all in or all out
    }
    store_run(r0 + rl, c0 + 4 * run, code);
}

// ---------------------------------------------------------------------------
// K2 WTALeft + WTARight of one frame's raw cost volume, costs recomputed from
// census (census.cpp:81-82: popcount of the XOR; pairs with d > c are the
// 255 entries and are never the minimum).
//   L(r,c) = argmin_{d <= min(c,D-1)} pop(CC[r][c] ^ CN[r][c-d])   (postprocess.cpp:74-185)
//   R(r,c) = argmin_{k <= min(D-1,C-1-c)} pop(CC[r][c+k] ^ CN[r][c]) (postprocess.cpp:187-315)
// Lowest index wins ties: keys are (cost << 9) | d, reduced with min.
// One block = one frame row segment of 256 pixels; codes staged in LDS.
// ---------------------------------------------------------------------------
// Rows of the deep padding bands other than a band's first row (band_row):
// the WTA kernels compute the first row of each band and k_wta_fill copies it.
__device__ __forceinline__ bool wta_band_copy(int r, int R, int D) { return band_row(r, R, D) != r; }

__global__ __launch_bounds__(256) void k_wta(const RigDesc* __restrict__ rigs, int nrig,
                                             const int* __restrict__ bs)
{
    extern __shared__ uint64_t lds64[];
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    int lb = blockIdx.x - bs[ri];
    const int D = g.D;
    const int n02 = g.Hp * ((g.Wp + 255) / 256), n13 = g.Wp * ((g.Hp + 255) / 256);
    int F;
    if (lb < n02) F = 0;
    else if ((lb -= n02) < n13) F = 1;
    else if ((lb -= n13) < n02) F = 2;
    else { lb -= n02; F = 3; }
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const int tilesC = (C + 255) / 256;
    const int r = lb / tilesC, c0 = (lb % tilesC) * 256;
    if (wta_band_copy(r, R, D)) return;   // copied from the band's first row
    const uint64_t* cc = g.code[2 * F] + (size_t)r * C;
    const uint64_t* cn = g.code[2 * F + 1] + (size_t)r * C;
    const int n = 256 + D - 1;
    uint64_t* cnS = lds64;        // cnS[i] = CN[c0 - D + 1 + i]
    uint64_t* ccS = lds64 + n;    // ccS[i] = CC[c0 + i]
    for (int i = threadIdx.x; i < n; i += 256) {
        int a = c0 - D + 1 + i;
        cnS[i] = (a >= 0 && a < C) ? cn[a] : 0;
        int b = c0 + i;
        ccS[i] = (b < C) ? cc[b] : 0;
    }
    __syncthreads();
    const int tid = threadIdx.x;
    const int c = c0 + tid;
    if (c >= C) return;
    const uint64_t myc = ccS[tid];
    const uint64_t myn = cnS[tid + D - 1];
    unsigned bestL = 0xFFFFFFFFu, bestR = 0xFFFFFFFFu;
    const int lastL = min(c, D - 1);
    for (int d = 0; d <= lastL; d++) {
        unsigned key = ((unsigned)__popcll(myc ^ cnS[tid + D - 1 - d]) << 9) | (unsigned)d;
        bestL = min(bestL, key);
    }
    const int lastR = min(D - 1, C - 1 - c);
    for (int k = 0; k <= lastR; k++) {
        unsigned key = ((unsigned)__popcll(ccS[tid + k] ^ myn) << 9) | (unsigned)k;
        bestR = min(bestR, key);
    }
    g.wta[2 * F][(size_t)r * C + c] = (uint16_t)(bestL & 511u);
    g.wta[2 * F + 1][(size_t)r * C + c] = (uint16_t)(bestR & 511u);
}

// WTALeft and WTARight of one frame row from shared costs (same results as
// k_wta).  The right map's candidates for pixel p are cost(p + k, k), the
// left map's for pixel c are cost(c, d): one cost serves both.  One block per
// frame row, the row's codes in LDS; a wave takes 256-pixel segments, lane l
// the 4 pixels c = x0 + 4l + i, and steps d = 0..D-1:
//   w[i]  = CN[c_i - d]        partner codes: shift by one pixel per step,
//                              one new LDS read per step for 4 costs
//   key   = (pop(CC[c_i] ^ w[i]) << 9) | d        (lowest index wins ties)
//   L[i]  = min(L[i], key)                         left map of pixel c_i
//   R[i]  = min(R[i], key)                         right map of pixel c_i - d
// kept as their complements ~key under max (one v_mad_i32_i24 forms ~key =
// ~d - 512 pop), so that 0 is the neutral accumulator.
// R[i] belongs to pixel x0 + 4l + i - d, so the right accumulators move one
// pixel per step: R[3] of lane l becomes R[0] of lane l + 1 (DPP wave_shr:1,
// whose zero fill is the fresh accumulator entering lane 0: one move).
// Accumulators leaving lane 63, and all of them after the last step, are
// partial results (their pixel's other candidates lie in neighbouring
// segments): they meet by LDS atomic max, once per pixel and segment.  About
// 3.6 VALU per evaluation against 6.
// ---------------------------------------------------------------------------
constexpr int kWtaRowMaxWaves = 16;

size_t wta_row_lds(int C, int D)
{
    return ((size_t)C + (size_t)C + D - 1 + 256) * sizeof(uint64_t) + (size_t)C * sizeof(uint32_t);
}

// LFIX: a segment with pixels c < D - 1, whose left candidates d > c do not
// exist: their keys are kept out of the left minima (the same costs feed the
// right accumulators of pixels c - d < 0, which are dropped).  RFIX: a segment
// reaching past the row: its pixels c >= C add 128 to their popcounts (the
// popcount's addend), so their keys exceed every real one in the right minima
// at no cost per step; their left results are not stored.
template <bool LFIX, bool RFIX>
__device__ __forceinline__ void wta_row_segment(const uint64_t* ccS, const uint64_t* cnS, uint32_t* RP,
                                                uint16_t* outL, int x0, int C, int D, int lane)
{
    const int cb = x0 + 4 * lane;
    uint64_t cci[4], w[4];
    uint32_t bL[4], bR[4], ba[4];
    const int32_t m512 = -512;   // a register: the mad's other operand is the scalar ~d
#pragma unroll
    for (int i = 0; i < 4; i++) {
        cci[i] = (!RFIX || cb + i < C) ? ccS[cb + i] : 0;
        ba[i] = (RFIX && cb + i >= C) ? 128u : 0u;
        w[i] = cnS[cb + i + D - 1];
        bL[i] = 0u;
        bR[i] = 0u;
    }
    auto step = [&](int d) __attribute__((always_inline)) {
        const uint32_t nd = ~(uint32_t)d;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint64_t x = cci[i] ^ w[i];
            const uint32_t pop = bcnt_add((uint32_t)(x >> 32), bcnt_add((uint32_t)x, ba[i]));
            uint32_t nk;   // ~((pop << 9) | d) = ~d - 512 pop (pop <= 192)
            asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(nk) : "v"(pop), "v"(m512), "s"(nd));
            bL[i] = max(bL[i], (LFIX && d > cb + i) ? 0u : nk);
            bR[i] = max(bR[i], nk);
        }
    };
    auto shift = [&](int d) __attribute__((always_inline)) {
        // to step d + 1: R accumulators move up one pixel (one lane per 4)
        const uint32_t out = bR[3];
        const uint32_t in = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)out, 0x138, 0xF, 0xF, true);
        const int pl = x0 + 255 - d;          // pixel of lane 63's leaving accumulator
        // interior segments: x0 + 256 <= C and x0 >= D - 1, so 0 <= pl < C; at
        // the row's start pl < 0 once d > x0 + 255 (D > 256)
        if (lane == 63 && (!LFIX || pl >= 0) && (!RFIX || pl < C)) atomicMax(&RP[pl], out);
        bR[3] = bR[2]; bR[2] = bR[1]; bR[1] = bR[0]; bR[0] = in;
        w[3] = w[2]; w[2] = w[1]; w[1] = w[0];
        w[0] = cnS[cb + D - 2 - d];           // CN[cb - (d + 1)]
    };
    // four steps per iteration (the rotation of w and R repeats every four
    // shifts, so the registers rename with no moves); D % 8 == 0, and the
    // last quad, which has no final shift, is peeled
    int d = 0;
    for (; d + 4 < D; d += 4) {
        step(d);     shift(d);
        step(d + 1); shift(d + 1);
        step(d + 2); shift(d + 2);
        step(d + 3); shift(d + 3);
    }
    step(d);     shift(d);
    step(d + 1); shift(d + 1);
    step(d + 2); shift(d + 2);
    step(d + 3);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (!RFIX || cb + i < C) G(outL)[cb + i] = (uint16_t)(~bL[i] & 511u);
        const int pp = cb + i - (D - 1);      // pixel of R[i] after the last step
        if (pp >= 0 && pp < C) atomicMax(&RP[pp], bR[i]);
    }
}

template <bool V>
__global__ __launch_bounds__(64 * kWtaRowMaxWaves) void k_wta_row(const RigDesc* __restrict__ rigs, int nrig,
                                                                 const int* __restrict__ bs)
{
    extern __shared__ uint64_t lds64[];
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lb = blockIdx.x - bs[ri];
    const int D = g.D;
    // V: frames 1 and 3 (W' rows of H' pixels), else frames 0 and 2 (H' rows of W')
    const int nrow = V ? g.Wp : g.Hp, C = V ? g.Hp : g.Wp;
    const int F = (V ? 1 : 0) + (lb >= nrow ? 2 : 0);
    const int r = lb >= nrow ? lb - nrow : lb;
    if (wta_band_copy(r, nrow, D)) return;   // copied from the band's first row
    const uint64_t* cc = g.code[2 * F] + (size_t)r * C;
    const uint64_t* cn = g.code[2 * F + 1] + (size_t)r * C;
    uint64_t* ccS = lds64;                       // ccS[c] = CC[c]
    uint64_t* cnS = lds64 + C;                   // cnS[i] = CN[i - (D - 1)], zero outside
    uint32_t* RP = (uint32_t*)(cnS + C + D - 1 + 256);   // right-map partial keys
    const int ncn = C + D - 1 + 256;
    for (int i = threadIdx.x; i < ncn; i += blockDim.x) {
        const int a = i - (D - 1);
        cnS[i] = (a >= 0 && a < C) ? G(cn)[a] : 0;
    }
    for (int i = threadIdx.x; i < C; i += blockDim.x) {
        ccS[i] = G(cc)[i];
        RP[i] = 0u;   // complemented keys under max
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
    const int nseg = (C + 255) >> 8;
    uint16_t* const outL = g.wta[2 * F] + (size_t)r * C;
    for (int sg = wave; sg < nseg; sg += nwaves) {
        const int x0 = sg * 256;
        // segments at the row's ends mask the candidates that do not exist
        // (fewer than D at the start, pixels past the row at the end)
        const bool lfix = x0 < D - 1, rfix = x0 + 256 > C;
        if (lfix && rfix) wta_row_segment<true, true>(ccS, cnS, RP, outL, x0, C, D, lane);
        else if (lfix) wta_row_segment<true, false>(ccS, cnS, RP, outL, x0, C, D, lane);
        else if (rfix) wta_row_segment<false, true>(ccS, cnS, RP, outL, x0, C, D, lane);
        else wta_row_segment<false, false>(ccS, cnS, RP, outL, x0, C, D, lane);
    }
    __syncthreads();
    uint16_t* const outR = g.wta[2 * F + 1] + (size_t)r * C;
    for (int p = threadIdx.x; p < C; p += blockDim.x) G(outR)[p] = (uint16_t)(~RP[p] & 511u);
}

// The WTA maps of the deep padding bands' rows from each band's first row:
// one block per (map, band, 1024 columns), 4 columns per thread.
constexpr int kWtaFillCols = 1024;
__global__ __launch_bounds__(256) void k_wta_fill(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lb = blockIdx.x - bs[ri];
    const int nch = (max(g.Wp, g.Hp) + kWtaFillCols - 1) / kWtaFillCols;
    const int m = lb / (2 * nch), band = (lb / nch) & 1;
    const int F = m >> 1;
    const int R = frame_rows(g, F), C = frame_cols(g, F), D = g.D;
    const int c = (lb % nch) * kWtaFillCols + 4 * (int)threadIdx.x;   // C % 4 == 0
    if (c >= C) return;
    const int lo = band ? R - D + 3 : 3, hi = band ? R - 4 : D - 4;
    uint16_t* const map = g.wta[m];
    const uint2 v = *(const uint2*)(map + (size_t)lo * C + c);
    for (int r = lo + 1; r <= hi; r++) *(uint2*)(map + (size_t)r * C + c) = v;
}

// compile-time loop: f(integral_constant<int, I>) for I in [B, E), so register
// arrays indexed by I are never indexed at run time
template <int I> struct IC { static constexpr int value = I; };
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(IC<B>{});
        static_for<B + 1, E>(f);
    }
}

// ---------------------------------------------------------------------------
// K3 in-place 3x3 median (postprocess.cpp:15-71 called with source == dest).
// Over flat index q of a R x C map, processed in increasing q:
//   out[q] = in[q]                       q < C  or  q >= RC-C-4
//   out[C] = 0                           (first store is the zero lastMedian)
//   out[q] = med9(out[q-C-1], out[q-C], out[q-C+1],
//                 in[q-1],  in[q],  in[q+1],  in[q+C-1], in[q+C], in[q+C+1])
//                                        C+1 <= q <= RC-C-5
// The six `in` values of each q are fixed, so a parallel prepass sorts them
// (s0 <= ... <= s5) and keeps s1..s4.  With the three recursive values sorted
// (u0 <= u1 <= u2), the 5th smallest of the nine is
//   min(s4, max(s3, u0), max(s2, u1), max(s1, u2))
// (k-th smallest of the union of two sorted lists; s0 and s5 can never be the
// median).  The serial pass is then a few min/max per pixel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cswap(unsigned& a, unsigned& b)
{
    unsigned lo = min(a, b), hi = max(a, b);
    a = lo;
    b = hi;
}

__global__ __launch_bounds__(256) void k_median_prep(const RigDesc* __restrict__ rigs, int nrig,
                                                     const int* __restrict__ bs, uint64_t* const* __restrict__ scratch)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const long npx = (long)g.Wp * g.Hp;
    const long perMap = (npx + 255) / 256;
    const long lb = blockIdx.x - bs[ri];
    const int m = (int)(lb / perMap);
    const long q = (lb % perMap) * 256 + threadIdx.x;
    const int F = m >> 1;
    const long C = frame_cols(g, F);
    if (q >= npx) return;
    const uint16_t* in = g.wta[m];
    uint64_t* out = scratch[ri] + (size_t)m * npx;
    if (q < C + 1 || q > npx - C - 5) { out[q] = 0; return; }
    unsigned v0 = in[q - 1], v1 = in[q], v2 = in[q + 1];
    unsigned v3 = in[q + C - 1], v4 = in[q + C], v5 = in[q + C + 1];
    // 12-comparator sorting network for six inputs
    cswap(v0, v5); cswap(v1, v3); cswap(v2, v4);
    cswap(v1, v2); cswap(v3, v4);
    cswap(v0, v3); cswap(v2, v5);
    cswap(v0, v1); cswap(v2, v3); cswap(v4, v5);
    cswap(v1, v2); cswap(v3, v4);
    out[q] = (uint64_t)v1 | ((uint64_t)v2 << 16) | ((uint64_t)v3 << 32) | ((uint64_t)v4 << 48);
}

__device__ __forceinline__ unsigned med9_4(unsigned s1, unsigned s2, unsigned s3, unsigned s4, unsigned a,
                                           unsigned b, unsigned c)
{
    cswap(a, b); cswap(b, c); cswap(a, b);
    return min(min(s4, max(s3, a)), min(max(s2, b), max(s1, c)));
}

__device__ __forceinline__ unsigned med9_pre(uint64_t s, unsigned a, unsigned b, unsigned c)
{
    return med9_4((unsigned)(s & 0xFFFF), (unsigned)((s >> 16) & 0xFFFF), (unsigned)((s >> 32) & 0xFFFF),
                  (unsigned)(s >> 48), a, b, c);
}

// the same from the six fixed values themselves (12-comparator network)
__device__ __forceinline__ unsigned med9_raw(const uint32_t (&w)[6], unsigned a, unsigned b, unsigned c)
{
    unsigned v0 = w[0], v1 = w[1], v2 = w[2], v3 = w[3], v4 = w[4], v5 = w[5];
    cswap(v0, v5); cswap(v1, v3); cswap(v2, v4);
    cswap(v1, v2); cswap(v3, v4);
    cswap(v0, v3); cswap(v2, v5);
    cswap(v0, v1); cswap(v2, v3); cswap(v4, v5);
    cswap(v1, v2); cswap(v3, v4);
    return med9_4(v1, v2, v3, v4, a, b, c);
}

template <int K, bool RAW> struct MedBuf;
template <int K> struct MedBuf<K, true> { uint32_t v[K][6]; uint32_t l[6]; };   // six fixed values
template <int K> struct MedBuf<K, false> { uint64_t v[K]; uint64_t l; };        // their s1..s4

// One block per (rig, map), 1024 threads; rows in sequence, columns in
// parallel: thread t owns columns t + 1024 k for k < K = ceil(C / 1024), and
// frame rows are at most kMedMaxK * 1024 wide.  The order statistics of row
// r + 1 are loaded while row r is computed (two register buffers,
// alternating), and each result is stored from the register that holds it.
// LDS keeps two rows of results, each with one slot in front (index -1)
// holding out(r - 2, C - 1), the raster-order predecessor of column 0 in the
// window of row r; so every column reads its three upper neighbours alike.
// The last column reads column 0 of the same row, which thread 0 computes
// itself, so one barrier per row suffices.  Rows 2 .. R - 3 have none of the
// special cases (out(1, 0) = 0 and the copied tail from index R*C - C - 4 on)
// and run without checks.  EXACT: K is the frame's exact slot count, so only
// slot K - 1 can reach past column C - 2.  PF: load row r + 1 during row r
// (frames up to 8192 wide); without it the loads of a row precede its
// compute, which keeps the 16-slot buffer of the widest frames in registers.
// RAW (frames up to 4096 wide): the six fixed values of each position are
// loaded a row ahead and sorted in the row's step, so no prepass runs; wider
// frames read the prepass's order statistics (fewer registers per slot).
constexpr int kMedMaxK = 16;
constexpr int kMedRawK = 4;
// row buffers of the serial median's prefetch ring: the sorting form holds six
// values per slot (three buffers fit its register budget, four spill), the
// prepass form one (four fit); frames wider than the sorting form's keep two
#ifndef SISTER_MED_RING_RAW
#define SISTER_MED_RING_RAW 3
#endif
#ifndef SISTER_MED_RING_PRE
#define SISTER_MED_RING_PRE 4
#endif
template <int K, bool RAW>
constexpr int med_ring() { return RAW ? SISTER_MED_RING_RAW : K <= kMedRawK ? SISTER_MED_RING_PRE : 2; }

template <int K, bool EXACT, bool PF, bool RAW>
__device__ __forceinline__ void median_serial_body(int R, int C, const uint16_t* __restrict__ in,
                                                   uint16_t* __restrict__ out,
                                                   const uint64_t* __restrict__ pre, uint16_t* lds)
{
    const int tid = threadIdx.x;
    const bool wave0 = __builtin_amdgcn_readfirstlane(tid) < 64;
    const int qhi = R * C - C - 5;
    const int kact = (C + 1023) >> 10;
    uint16_t* prev = lds + 1;
    uint16_t* cur = lds + C + 2;
    for (int c = tid; c < C; c += 1024) {
        const uint16_t v = G(in)[c];
        prev[c] = v;
        G(out)[c] = v;
    }
    if (tid == 0) prev[-1] = 0;
    // slot k may hold a column below C - 1 (one that this thread computes)
    auto live = [&](int k) __attribute__((always_inline)) {
        return (EXACT && k < K - 1) || tid + 1024 * k < C - 1;
    };
    auto col = [&](int k) __attribute__((always_inline)) {
        return (EXACT && k < K - 1) ? tid + 1024 * k : min(tid + 1024 * k, C - 2);
    };
    // Order statistics of row r (1 <= r < R).  Indices past qhi (the copied
    // tail) are clamped; their statistics are loaded but never used.  Byte
    // offsets fit 32 bits (R * C * 8 <= 2^31).  The last column's statistics
    // are needed by thread 0 alone; wave 0 loads them with a lane-dependent
    // index (lane 0's is the one used) so that the load stays a vector load:
    // a scalar load would make every LDS wait of the row wait for it too.
    auto at = [&](int i) __attribute__((always_inline)) {
        return G((const uint64_t*)((const char*)pre + (uint32_t)i * 8u));
    };
    // RAW: in[q-1..q+1] and in[q+C-1..q+C+1] at q clamped to qhi (C <= q:
    // every index in bounds; the clamped positions' values go unused)
    auto six = [&](int q, uint32_t (&o)[6]) __attribute__((always_inline)) {
        const int qq = min(q, qhi);
        o[0] = G(in)[qq - 1]; o[1] = G(in)[qq]; o[2] = G(in)[qq + 1];
        o[3] = G(in)[qq + C - 1]; o[4] = G(in)[qq + C]; o[5] = G(in)[qq + C + 1];
    };
    auto fetch = [&](int r, MedBuf<K, RAW>& b) __attribute__((always_inline)) {
        const int base = r * C;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!EXACT && k >= kact) break;
            if constexpr (RAW) six(base + col(k), b.v[k]);
            else b.v[k] = *at(min(base + col(k), qhi));
        }
        if constexpr (RAW) {
            if (wave0) six(base + C - 1 + tid, b.l);
        } else {
            if (wave0) b.l = *at(min(base + C - 1 + tid, qhi));
        }
    };
    auto med = [&](const MedBuf<K, RAW>& b, int k, unsigned a, unsigned bb, unsigned c) __attribute__((always_inline)) {
        if constexpr (RAW) return k < 0 ? med9_raw(b.l, a, bb, c) : med9_raw(b.v[k], a, bb, c);
        else return k < 0 ? med9_pre(b.l, a, bb, c) : med9_pre(b.v[k], a, bb, c);
    };
    // In rows 2 .. R - 3 nothing is read from `in`; the other rows wait for
    // their `in` loads inside their own branch (vmcnt(0), expcnt and lgkmcnt
    // left free: 0x0F70), so that the plain rows' stores do not wait for the
    // next row's statistics still in flight.
    auto row = [&](int r, const MedBuf<K, RAW>& pb) __attribute__((always_inline)) {
        const int base = r * C;
        const bool plain = r >= 2 && r <= R - 3;
        unsigned v[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!EXACT && k >= kact) break;
            const int c = col(k);
            v[k] = med(pb, k, prev[c - 1], prev[c], prev[c + 1]);
        }
        if (!plain) {
            // out(1, 0) = 0; from index qhi + 1 on the input is copied
#pragma unroll
            for (int k = 0; k < K; k++) {
                if (!EXACT && k >= kact) break;
                const int q = base + tid + 1024 * k;
                if (q == C) v[k] = 0;
                else if (q > qhi && live(k)) v[k] = G(in)[q];
            }
            __builtin_amdgcn_s_waitcnt(0x0F70);
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (!EXACT && k >= kact) break;
            if (live(k)) {
                const int c = tid + 1024 * k;
                cur[c] = (uint16_t)v[k];
                G(out)[base + c] = (uint16_t)v[k];
            }
        }
        if (tid == 0) {
            // column C - 1 reads column 0 of this row, v[0] of this thread
            const int q = base + C - 1;
            unsigned w = med(pb, -1, prev[C - 2], prev[C - 1], v[0]);
            if (!plain) {
                if (q > qhi) w = G(in)[q];
                __builtin_amdgcn_s_waitcnt(0x0F70);
            }
            cur[C - 1] = (uint16_t)w;
            cur[-1] = prev[C - 1];
            G(out)[q] = (uint16_t)w;
        }
        __syncthreads();
        uint16_t* t = prev; prev = cur; cur = t;
    };
    if constexpr (PF) {
        // a ring of kMedRing buffers: row r's statistics are requested
        // kMedRing - 1 rows ahead of their use, so that the row chain waits
        // on load latency only when a row computes faster than that
        constexpr int NB = med_ring<K, RAW>();
        MedBuf<K, RAW> buf[NB];
        static_for<0, NB - 1>([&](auto ic) { fetch(min(1 + decltype(ic)::value, R - 1), buf[decltype(ic)::value]); });
        __syncthreads();
        int r = 1;
        for (; r + NB - 1 < R; r += NB) {
            static_for<0, NB>([&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                fetch(min(r + i + NB - 1, R - 1), buf[(i + NB - 1) % NB]);
                row(r + i, buf[i]);
            });
        }
        // fewer than NB rows remain; all of them are fetched already
        static_for<0, NB - 1>([&](auto ic) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            if (r + i < R) row(r + i, buf[i]);
        });
    } else {
        MedBuf<K, RAW> buf;
        __syncthreads();
        for (int r = 1; r < R; r++) {
            fetch(r, buf);
            row(r, buf);
        }
    }
}

// MODE 0: frames up to kMedRawK * 1024 wide, sorting the six fixed values
// itself (its blocks of wider frames return at once); MODE 1: the same frames
// from the prepass; MODE 2: every frame from the prepass, for plans with
// frames wider than that (all maps at once: a rig's wide and narrow frames
// take the same time, so they run side by side).
// Blocks are dispatched long maps first: block b takes map slot b / nrig of rig
// b % nrig, slots 0-3 being the rig's four maps of the frames with more rows
// (the serial chain is one row per step) and slots 4-7 the other four.  When
// the blocks take more than one round of the CUs, the short maps then fill
// the CUs the long ones leave, instead of long maps starting last.
template <int MODE>
__global__ __launch_bounds__(1024) void k_median_serial(const RigDesc* __restrict__ rigs, int nrig,
                                                        uint64_t* const* __restrict__ scratch)
{
    extern __shared__ uint16_t lds16[];
    const int ri = blockIdx.x % nrig, slot = blockIdx.x / nrig;
    const RigDesc& g = rigs[ri];
    // frames 1/3 have W' rows, frames 0/2 H' rows; map m = 2 F + side
    const int long_odd = g.Wp > g.Hp ? 1 : 0;
    const int fr = (slot >> 1) & 1, side = slot & 1;             // frame pair index, side
    const int F = ((slot < 4) == (long_odd == 1)) ? 2 * fr + 1 : 2 * fr;
    const int m = 2 * F + side;
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const uint16_t* in = g.wta[m];
    uint16_t* out = g.med[m];
    const uint64_t* pre = scratch[ri] + (size_t)m * R * C;
    const int k = (C + 1023) >> 10;
    if (MODE < 2 && k > kMedRawK) return;
    constexpr bool RAW = MODE == 0;
    if (k <= kMedRawK) {
        switch (k) {
        case 1: median_serial_body<1, true, true, RAW>(R, C, in, out, pre, lds16); break;
        case 2: median_serial_body<2, true, true, RAW>(R, C, in, out, pre, lds16); break;
        case 3: median_serial_body<3, true, true, RAW>(R, C, in, out, pre, lds16); break;
        default: median_serial_body<4, true, true, RAW>(R, C, in, out, pre, lds16); break;
        }
    } else if constexpr (MODE == 2) {
        if (k <= 8) median_serial_body<8, false, true, false>(R, C, in, out, pre, lds16);
        else median_serial_body<kMedMaxK, false, false, false>(R, C, in, out, pre, lds16);
    }
}

// ---------------------------------------------------------------------------
// K4 LR check + mask (postprocess.cpp:318-341, hpp:201-206 etc.): frame pixel
// (r,c) keeps L when 0 <= L <= c and |L - R[c-L]| <= 5; mask = kept && L > 0
// && c >= D, written at the padded original coordinates (hpp:216,235-236,250-251).
// ---------------------------------------------------------------------------
// One block per 32x32 tile of frame pixels (4 per thread), read along frame
// rows.  Frames 0/180 store along image rows directly; frames 90/270, whose
// frame rows are image columns, pass the tile through LDS so that their
// stores also run along image rows.
constexpr int kMT32 = 32;

__global__ __launch_bounds__(256) void k_mask(const RigDesc* __restrict__ rigs, int nrig,
                                              const int* __restrict__ bs, int fmask)
{
    __shared__ uint8_t tile[kMT32][kMT32 + 1];
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int T = ((g.Wp + kMT32 - 1) / kMT32) * ((g.Hp + kMT32 - 1) / kMT32);
    const int lb = blockIdx.x - bs[ri];
    const int F = lb / T, t = lb % T;
    if (!(fmask & (1 << F))) return;
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const int tilesC = (C + kMT32 - 1) / kMT32;
    const int r0 = (t / tilesC) * kMT32, c0 = (t % tilesC) * kMT32;
    const int tc = threadIdx.x % kMT32, tr = threadIdx.x / kMT32;
    const uint16_t* ml = g.med[2 * F];
    const uint16_t* mr = g.med[2 * F + 1];
    uint8_t* const mk = g.mask[F];
    const bool odd = F & 1;
#pragma unroll
    for (int k = 0; k < kMT32 / 8; k++) {
        const int r = r0 + tr + 8 * k, c = c0 + tc;
        if (r >= R || c >= C) continue;
        const size_t q = (size_t)r * C + c;
        const int L = G(ml)[q];
        unsigned ok = 0;
        if (c >= g.D && L >= 1 && L <= c) {
            const int Rv = G(mr)[q - L];
            ok = abs(L - Rv) <= kLrThreshold;
        }
        if (odd) {
            tile[tr + 8 * k][tc] = (uint8_t)ok;
        } else {
            int y, x;
            frame_to_orig(g, F, r, c, y, x);
            G(mk)[(size_t)y * g.Wp + x] = (uint8_t)ok;
        }
    }
    if (!odd) return;
    __syncthreads();
    // frame (r, c) -> image (y, x): x = W'-1-r for both frames, y = H'-1-c
    // (90) or c (270); a thread takes frame row r0 + tc (image column) and
    // columns c0 + tr + 8 k (image rows), so lanes store along image rows
#pragma unroll
    for (int k = 0; k < kMT32 / 8; k++) {
        const int r = r0 + tc, c = c0 + tr + 8 * k;
        if (r >= R || c >= C) continue;
        int y, x;
        frame_to_orig(g, F, r, c, y, x);
        G(mk)[(size_t)y * g.Wp + x] = tile[tc][tr + 8 * k];
    }
}

// ---------------------------------------------------------------------------
// K5 mask-weighted cost sum (hpp:255-277).  M(y,x,d) = sum over used frames F
// with mask_F(y,x) = 1 of pop(CC_F[r][c] ^ CN_F[r][c-d]), (r,c) the frame
// pixel of (y,x).  A set mask implies c >= D > d, so every term is a plain
// popcount (<= 63) and M <= 252 fits a byte.
// Each frame's partners of a pixel lie along its frame row, which is an
// image row for frames 0/180 and an image column for frames 90/270, so two
// launches share the work: k_msum_run<0> (frames 0 and 180, one block = a run
// of kRun pixels of one row) writes M, k_msum_run<1> (frames 90 and 270, one
// block = a run of one column) adds to it.  Each block stages the partner
// window of its run (kRun + D - 1 codes per frame) in LDS once.  A wave takes
// one pixel at a time, lane q the disparities q + 64 j: conflict-free 8-byte
// LDS reads at immediate offsets and coalesced 64-byte stores per group, with
// no per-disparity branches.  Cell stride Dp, zero past D.
// ---------------------------------------------------------------------------
constexpr int kRun = 256;   // pixels per block
constexpr int kWinPad = 8;  // window entries before each partner window

// acc[j] += pop(cc ^ window[base - d]) for lane q's disparities d = q + 64 j.
// In the last group, lanes at or past Dp read as lane qlast = Dp - 1 - 64 (J-1)
// (their sums are not stored) and lanes at or past D are zeroed by the caller.
template <int J>
__device__ __forceinline__ void msum_add_frame(const uint64_t* L, uint64_t cc, int base, int q, int qlast,
                                               unsigned (&acc)[J])
{
    uint64_t tv[J];
    const uint64_t* w = L + base - q;       // w[-64 j] = window[base - d]
#pragma unroll
    for (int j = 0; j < J - 1; j++) tv[j] = w[-64 * j];
    tv[J - 1] = L[base - qlast - 64 * (J - 1)];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint64_t x = cc ^ tv[j];
        acc[j] = bcnt_add((uint32_t)(x >> 32), bcnt_add((uint32_t)x, acc[j]));
    }
}

// SIDE = 0: frames 0/180 along row y; SIDE = 1: frames 90/270 along column x.
// J = ceil(Dp / 64) groups of 64 disparities; every rig of the launch has this Dp.
template <int SIDE, int J>
__global__ __launch_bounds__(256) void k_msum_run(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs, int fuse)
{
    extern __shared__ uint64_t lds64[];
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lb = blockIdx.x - bs[ri];
    const int Wp = g.Wp, Hp = g.Hp, D = g.D, Dp = g.Dp;
    const int FA = SIDE == 0 ? 0 : 1, FB = SIDE == 0 ? 2 : 3;   // frame pair
    const bool useA = (fuse >> FA) & 1, useB = (fuse >> FB) & 1;
    // the run: SIDE 0 -> row y, x in [p0, p0+kRun); SIDE 1 -> column x, y in [p0, p0+kRun)
    const int runs = SIDE == 0 ? (Wp + kRun - 1) / kRun : (Hp + kRun - 1) / kRun;
    const int line = lb / runs, p0 = (lb % runs) * kRun;
    const int len = SIDE == 0 ? Wp : Hp;        // pixels along the line
    const int SPAN = kRun + D - 1;
    // lanes of the last group past D read up to Dp - D <= 8 entries before
    // their window; those terms are masked
    uint64_t* LA = lds64 + kWinPad;             // partners of frame FA
    uint64_t* LB = LA + SPAN + kWinPad;         // partners of frame FB
    uint64_t* LCa = LB + SPAN;                  // centers
    uint64_t* LCb = LCa + kRun;
    uint8_t* LM = (uint8_t*)(LCb + kRun);
    const int tid = threadIdx.x;
    // frame row holding the line and the frame column of run pixel i:
    //  SIDE 0: F0 row y col x;  F2 row y col W'-1-x
    //  SIDE 1: F1 row W'-1-x col H'-1-y;  F3 row W'-1-x col y
    const int frow = band_row(SIDE == 0 ? line : Wp - 1 - line, SIDE == 0 ? Hp : Wp, D);   // stored row
    const int fcols = SIDE == 0 ? Wp : Hp;
    const uint64_t* NA = g.code[2 * FA + 1] + (size_t)frow * fcols;
    const uint64_t* NB = g.code[2 * FB + 1] + (size_t)frow * fcols;
    const uint64_t* CA = g.code[2 * FA] + (size_t)frow * fcols;
    const uint64_t* CB = g.code[2 * FB] + (size_t)frow * fcols;
    // window of frame X: the frame columns [lo, lo + SPAN) covering every
    // partner col - d of the run (frame A columns run with the line
    // coordinate for F0 and F3, against it for F1 and F2)
    int loA, loB;
    if (SIDE == 0) { loA = p0 - (D - 1); loB = Wp - p0 - kRun - (D - 1); }
    else { loA = Hp - p0 - kRun - (D - 1); loB = p0 - (D - 1); }
    for (int k = tid - kWinPad; k < SPAN; k += 256) {
        const int ca = loA + k, cb = loB + k;
        LA[k] = (useA && k >= 0 && ca >= 0 && ca < fcols) ? G(NA)[ca] : 0;
        LB[k] = (useB && k >= 0 && cb >= 0 && cb < fcols) ? G(NB)[cb] : 0;
    }
    for (int i = tid; i < kRun; i += 256) {
        const int pos = p0 + i;                 // x (SIDE 0) or y (SIDE 1)
        uint64_t a = 0, b = 0;
        uint8_t mb = 0;
        if (pos < len) {
            const int colA = SIDE == 0 ? pos : Hp - 1 - pos;
            const int colB = SIDE == 0 ? Wp - 1 - pos : pos;
            const int y = SIDE == 0 ? line : pos, x = SIDE == 0 ? pos : line;
            const size_t q = (size_t)y * Wp + x;
            if (useA) a = G(CA)[colA];
            if (useB) b = G(CB)[colB];
            // a set mask implies frame column >= D; the check keeps indices in range
            if (useA && G(g.mask[FA])[q] && colA >= D) mb |= 1;
            if (useB && G(g.mask[FB])[q] && colB >= D) mb |= 2;
        }
        LCa[i] = a;
        LCb[i] = b;
        LM[i] = mb;
    }
    __syncthreads();
    // one wave per pixel; lane q owns d = q + 64 j.  In the last group, lanes
    // at or past Dp read as lane Dp - 1 - 64 (J - 1) (and store nothing),
    // lanes at or past D add nothing.
    const int wave = tid >> 6, q = tid & 63;
    const int qlast = min(q, Dp - 1 - 64 * (J - 1));
    const bool store_last = q + 64 * (J - 1) < Dp;
    const bool live_last = q + 64 * (J - 1) < D;
    uint8_t* const msum = g.msum;
    for (int i = wave; i < kRun; i += 4) {
        const int pos = p0 + i;
        if (pos >= len) break;
        const int y = SIDE == 0 ? line : pos, x = SIDE == 0 ? pos : line;
        gptr<uint8_t> o = G(msum) + ((size_t)y * Wp + x) * Dp;
        unsigned acc[J];
#pragma unroll
        for (int j = 0; j < J; j++) acc[j] = 0;
        if (SIDE == 1 && (fuse & 5)) {
#pragma unroll
            for (int j = 0; j < J - 1; j++) acc[j] = o[q + 64 * j];
            if (store_last) acc[J - 1] = o[q + 64 * (J - 1)];
        }
        const int mb = __builtin_amdgcn_readfirstlane(LM[i]);   // one pixel per wave
        if (mb & 1) msum_add_frame<J>(LA, LCa[i], (SIDE == 0 ? pos : Hp - 1 - pos) - loA, q, qlast, acc);
        if (mb & 2) msum_add_frame<J>(LB, LCb[i], (SIDE == 0 ? Wp - 1 - pos : pos) - loB, q, qlast, acc);
#pragma unroll
        for (int j = 0; j < J - 1; j++) o[q + 64 * j] = (uint8_t)acc[j];
        if (store_last) o[q + 64 * (J - 1)] = live_last ? (uint8_t)acc[J - 1] : (uint8_t)0;
    }
}

// Fused variant for cost sums over all four frames: a kMT x kMT tile of pixels
// needs the partner windows of every frame line through it (tile rows for
// frames 0/180, tile columns for frames 90/270: 4 kMT lines of kMT + D - 1
// codes) and 4 kMT^2 centre codes, staged in LDS, and writes M once -- no
// read-modify-write of M between two launches.  That is 118.5 KB of LDS at
// D = 192, so one block per CU: the blocks are persistent and load the next
// tile's windows into registers while the current tile is summed.  Block b
// takes tiles s(b), s(b) + n, s(b) + 2n, ... (n blocks); s groups the blocks
// that share an XCD (b mod 8 under round-robin dispatch) on consecutive
// tiles, so neighbouring tiles, whose row windows overlap, share an L2.  The
// two-launch form covers the disparity counts beyond what one block can hold.
// Per pixel as above: one wave per pixel, lane q owns d = q + 64 j.
// ---------------------------------------------------------------------------
constexpr int kMT = 16;
#ifndef SISTER_MSUM_HOIST
#define SISTER_MSUM_HOIST 0
#endif

// window line stride (entries) for J groups of 64 disparities: holds
// kWinPad + kMT + D - 1 <= 64 J + 23 entries; a compile-time constant, so a
// pixel's reads sit at immediate offsets from per-frame lane addresses
template <int J>
constexpr int msum_lstr() { return 64 * J + 24; }

size_t msum_tile_lds(int Dp)
{
    const int J = (Dp + 63) / 64;
    return ((size_t)4 * kMT * (64 * J + 24) + 4 * kMT * kMT) * 8 + kMT * kMT;
}

template <int J>
__global__ __launch_bounds__(1024) void k_msum_tile(const RigDesc* __restrict__ rigs, int nrig,
                                                    const int* __restrict__ bs, int ntiles, int fuse)
{
    static_assert(kMT == 16, "one wave per line of each frame, one thread per tile pixel below 256");
    // a line holds kWinPad + kMT + D - 1 <= 64 J + 23 entries: J + 1 chunks of 64
    constexpr int NK = J + 1;
    extern __shared__ uint64_t lds64[];
    const int D = rigs[0].D, Dp = rigs[0].Dp;   // one disparity count per launch
    const int SPAN = kMT + D - 1;
    constexpr int LSTR = msum_lstr<J>();
    auto win = [&](int F, int l) { return lds64 + (F * kMT + l) * LSTR + kWinPad; };
    uint64_t* CEN = lds64 + 4 * kMT * LSTR;     // [frame][pixel]
    uint8_t* MB = (uint8_t*)(CEN + 4 * kMT * kMT);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int nb = gridDim.x, b = blockIdx.x;
    int t = (nb & 7) == 0 ? (b & 7) * (nb >> 3) + (b >> 3) : b;

    struct Geo { int ri, x0, y0; };
    auto geo = [&](int tt) __attribute__((always_inline)) {
        Geo q;
        q.ri = find_rig(bs, nrig, tt);
        const int ntx = (rigs[q.ri].Wp + kMT - 1) / kMT, lb = tt - bs[q.ri];
        q.x0 = (lb % ntx) * kMT;
        q.y0 = (lb / ntx) * kMT;
        return q;
    };
    // Frame F pixel of image pixel (y, x) and the frame column of window
    // entry 0 of the line through the tile:
    //   F0: row y,        col x,         window from x0 - (D-1)
    //   F2: row y,        col W'-1-x,    window from W' - x0 - kMT - (D-1)
    //   F1: row W'-1-x,   col H'-1-y,    window from H' - y0 - kMT - (D-1)
    //   F3: row W'-1-x,   col y,         window from y0 - (D-1)
    // pc: centre codes; pk: mask bytes, read only at stash time, so that
    // nothing waits for the loads while the current tile is summed
    uint64_t pw[4][NK], pc[4];
    uint32_t pk[4];
    auto load = [&](const Geo& q) __attribute__((always_inline)) {
        const RigDesc& g = rigs[q.ri];
        const int Wp = g.Wp, Hp = g.Hp;
#pragma unroll
        for (int F = 0; F < 4; F++) {
            const int l = wave;
            const bool rowf = (F & 1) == 0;
            const bool line_ok = rowf ? q.y0 + l < Hp : q.x0 + l < Wp;
            const int frow = band_row(rowf ? q.y0 + l : Wp - 1 - (q.x0 + l), rowf ? Hp : Wp, D);   // stored row
            const int fcols = rowf ? Wp : Hp;
            const int lo = F == 0 ? q.x0 - (D - 1)
                         : F == 2 ? Wp - q.x0 - kMT - (D - 1)
                         : F == 1 ? Hp - q.y0 - kMT - (D - 1) : q.y0 - (D - 1);
            const bool use = ((fuse >> F) & 1) && line_ok;
            const uint64_t* N = g.code[2 * F + 1] + (size_t)(use ? frow : 0) * fcols;
#pragma unroll
            for (int kk = 0; kk < NK; kk++) {
                const int k = lane - kWinPad + 64 * kk, col = lo + k;
                pw[F][kk] = (use && k >= 0 && k < SPAN && col >= 0 && col < fcols) ? G(N)[col] : 0;
            }
        }
        if (tid < kMT * kMT) {
            const int y = q.y0 + tid / kMT, x = q.x0 + tid % kMT;
            const bool in = y < Hp && x < Wp;
#pragma unroll
            for (int F = 0; F < 4; F++) {
                const bool rowf = (F & 1) == 0;
                const int frow = band_row(rowf ? y : Wp - 1 - x, rowf ? Hp : Wp, D);   // stored row
                const int fcol = F == 0 ? x : F == 2 ? Wp - 1 - x : F == 1 ? Hp - 1 - y : y;
                const int fcols = rowf ? Wp : Hp;
                // a set mask implies frame column >= D; the check keeps indices in range
                const bool use = in && ((fuse >> F) & 1) && fcol >= D;
                pc[F] = use ? G(g.code[2 * F])[(size_t)frow * fcols + fcol] : 0;
                pk[F] = use ? G(g.mask[F])[(size_t)y * Wp + x] : 0;
            }
        }
    };
    auto stash = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int F = 0; F < 4; F++) {
            uint64_t* W = win(F, wave);
#pragma unroll
            for (int kk = 0; kk < NK; kk++) {
                const int k = lane - kWinPad + 64 * kk;
                if (k < SPAN) W[k] = pw[F][kk];
            }
        }
        if (tid < kMT * kMT) {
            uint32_t pm = 0;
#pragma unroll
            for (int F = 0; F < 4; F++) {
                CEN[F * kMT * kMT + tid] = pc[F];
                pm |= (pk[F] != 0 ? 1u : 0u) << F;
            }
            MB[tid] = (uint8_t)pm;
        }
    };
    const int q = lane;
    const int qlast = min(q, Dp - 1 - 64 * (J - 1));
    const bool store_last = q + 64 * (J - 1) < Dp;
    const bool live_last = q + 64 * (J - 1) < D;
    // wave w walks tile column w, its 16 pixels fully unrolled: every read
    // sits at a compile-time offset from one lane address per frame, so no
    // address arithmetic runs per pixel.  Frame line and window index of
    // pixel (ty, tx) at disparity d:
    //   F0: line ty, index tx + D - 1 - d;         F2: line ty, index (kMT-1-tx) + D - 1 - d
    //   F1: line tx, index (kMT-1-ty) + D - 1 - d; F3: line tx, index ty + D - 1 - d
    auto sum_tile = [&](const Geo& gq) __attribute__((always_inline)) {
        const RigDesc& g = rigs[gq.ri];
        const int Wp = g.Wp, Hp = g.Hp;
        const int tx = wave, x = gq.x0 + tx;
        if (x >= Wp) return;
        const uint32_t mcol = lane < kMT ? MB[lane * kMT + tx] : 0u;   // lane ty: mask bits of (ty, tx)
        // group j of lane q reads entry base - q - 64 j: groups below J - 1 from
        // pa + 64 (J - 1 - j), the last group (lanes past Dp read as qlast) from pl
        const int lo = D - 1 - 64 * (J - 1);
        const uint64_t* pa[4] = {win(0, 0) + tx + lo - q, win(1, tx) + lo - q,
                                 win(2, 0) + (kMT - 1 - tx) + lo - q, win(3, tx) + lo - q};
        const uint64_t* pl[4] = {win(0, 0) + tx + lo - qlast, win(1, tx) + lo - qlast,
                                 win(2, 0) + (kMT - 1 - tx) + lo - qlast, win(3, tx) + lo - qlast};
        const uint64_t* pc0 = CEN + tx;
        gptr<uint8_t> o = G(g.msum) + ((size_t)gq.y0 * Wp + x) * Dp;
        const size_t rstride = (size_t)Wp * Dp;
        static_for<0, kMT>([&](auto tyc) __attribute__((always_inline)) {
            constexpr int ty = decltype(tyc)::value;
            if (gq.y0 + ty >= Hp) return;
            unsigned acc[J];
#pragma unroll
            for (int j = 0; j < J; j++) acc[j] = 0;
            const int mb = __builtin_amdgcn_readlane(mcol, ty);   // one pixel per wave
#if SISTER_MSUM_HOIST
            // every frame's window and centre reads are issued before any
            // frame's (mask-dependent) sum, so one LDS latency covers the
            // pixel instead of one per frame
            uint64_t tv[4][J], cc[4];
            static_for<0, 4>([&](auto kc) __attribute__((always_inline)) {
                constexpr int F = decltype(kc)::value;
                constexpr int off = F == 0 || F == 2 ? ty * LSTR : F == 1 ? kMT - 1 - ty : ty;
#pragma unroll
                for (int j = 0; j < J - 1; j++) tv[F][j] = pa[F][off + 64 * (J - 1 - j)];
                tv[F][J - 1] = pl[F][off];
                cc[F] = pc0[F * kMT * kMT + ty * kMT];
            });
            static_for<0, 4>([&](auto kc) __attribute__((always_inline)) {
                constexpr int F = (0x3120 >> (4 * decltype(kc)::value)) & 15;   // frames 0, 2, 1, 3
                if (!(mb & (1 << F))) return;
#pragma unroll
                for (int j = 0; j < J; j++) {
                    const uint64_t xv = cc[F] ^ tv[F][j];
                    acc[j] = bcnt_add((uint32_t)(xv >> 32), bcnt_add((uint32_t)xv, acc[j]));
                }
            });
#else
            static_for<0, 4>([&](auto kc) __attribute__((always_inline)) {
                constexpr int F = (0x3120 >> (4 * decltype(kc)::value)) & 15;   // frames 0, 2, 1, 3
                constexpr int off = F == 0 || F == 2 ? ty * LSTR : F == 1 ? kMT - 1 - ty : ty;
                if (!(mb & (1 << F))) return;
                uint64_t tv[J];
#pragma unroll
                for (int j = 0; j < J - 1; j++) tv[j] = pa[F][off + 64 * (J - 1 - j)];
                tv[J - 1] = pl[F][off];
                const uint64_t cc = pc0[F * kMT * kMT + ty * kMT];
#pragma unroll
                for (int j = 0; j < J; j++) {
                    const uint64_t xv = cc ^ tv[j];
                    acc[j] = bcnt_add((uint32_t)(xv >> 32), bcnt_add((uint32_t)xv, acc[j]));
                }
            });
#endif
            gptr<uint8_t> op = o + ty * rstride;
#pragma unroll
            for (int j = 0; j < J - 1; j++) op[q + 64 * j] = (uint8_t)acc[j];
            if (store_last) op[q + 64 * (J - 1)] = live_last ? (uint8_t)acc[J - 1] : (uint8_t)0;
        });
    };
    Geo cur = geo(t);
    load(cur);
    for (;;) {
        stash();
        __syncthreads();
        const int tn = t + nb;
        Geo nxt = cur;
        if (tn < ntiles) {
            nxt = geo(tn);
            load(nxt);          // in flight while this tile is summed
        }
        sum_tile(cur);
        if (tn >= ntiles) break;
        __syncthreads();        // LDS free for the next stash
        t = tn;
        cur = nxt;
    }
}

// ---------------------------------------------------------------------------
// K6 SGM (sgm.cpp:26-455).  S is the sum of eight path costs; each path is a
// family of independent chains (rows, columns, diagonals, anti-diagonals), each
// swept in both directions.  One lane owns one chain and keeps all D
// disparities of its path state in registers as packed u16 pairs, so d-1 / d+1
// are a byte-permute of neighbouring registers, the minimum over d is a
// running packed min, and no lane talks to another.
//
// Standard step (rows >= 1 of a pass, all paths; sgm.cpp:278-383):
//   L(d) = M(d) + min(L'(d), min(L'(d-1), L'(d+1)) + P1, minL' + P2) - minL'
// with L'(-1) = L'(D) = guard.  (L - M) is in [0, P2] and M <= 252, so the
// reference's saturating u16 arithmetic never saturates and u16 lanes hold it.
// Pass structure reproduced exactly (sgm.cpp:89-207):
//   - on the first line of each pass (row 0 down, row H'-1 up) paths r1, r2,
//     r3 only take the state c~ = (M == 255 ? 0 : M) and add nothing to S; the
//     horizontal path r0 runs on c~ truncated to 8 bits (uint8 saturate_cast)
//     and is the only contribution of that pass on that row;
//   - r0 restarts every other row from L' = 0, minL' = 0 (L = M);
//   - r1 / r3 entering from beyond the left/right border see L' = guard and
//     minL' = 0, so L = M + P2 (the standard step from a guard state).
// Each family stores, for rows 1..H'-2, (L_fwd - M) + (L_bwd - M) in [0, 200]
// as a byte (part[f]; the rows family its two directions as separate bytes,
// part[0] and rowsb, so they run independently), and for rows 0 and H'-1 its
// exact contribution as u16 (edge[f]).  The anti-diagonal family runs last:
// its upward sweep forms
//   S = 8 M + part[0] + rowsb + part[1] + part[2] + its own (L - M) sum
// (rows 1..H'-2) and takes the winner (hpp:283) for cells inside the crop.
// ---------------------------------------------------------------------------
typedef unsigned short us2 __attribute__((ext_vector_type(2)));
constexpr uint32_t kGuard2 = 0x3FFF3FFFu;

__device__ __forceinline__ us2 U2(uint32_t v) { return __builtin_bit_cast(us2, v); }
__device__ __forceinline__ uint32_t W2(us2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ us2 umin2(us2 a, us2 b) { return __builtin_elementwise_min(a, b); }
// bytes b0,b1 of w -> pair (b0, b1); bytes b2,b3 -> pair (b2, b3)
__device__ __forceinline__ uint32_t bytes01(uint32_t w) { return __builtin_amdgcn_perm(w, w, 0x0c010c00u); }
__device__ __forceinline__ uint32_t bytes23(uint32_t w) { return __builtin_amdgcn_perm(w, w, 0x0c030c02u); }
// low bytes of pairs a, b -> one word (a.lo, a.hi, b.lo, b.hi)
__device__ __forceinline__ uint32_t pack4(uint32_t a, uint32_t b) { return __builtin_amdgcn_perm(b, a, 0x06040200u); }
__device__ __forceinline__ uint32_t dup16(uint32_t v) { return v | (v << 16); }
// 255 -> 0 in both halves (paramInvalidDispCost, sgm.cpp:109-110,146-147)
__device__ __forceinline__ uint32_t tilde2(uint32_t p)
{
    uint32_t lo = (p & 0xFFFFu) == 255u ? 0u : (p & 0xFFFFu);
    uint32_t hi = (p >> 16) == 255u ? 0u : (p >> 16);
    return lo | (hi << 16);
}

// unpack one 16-byte chunk of a u8 volume into 8 pairs, and back
__device__ __forceinline__ void unpack16(const uint4& v, uint32_t (&p)[8])
{
    p[0] = bytes01(v.x); p[1] = bytes23(v.x);
    p[2] = bytes01(v.y); p[3] = bytes23(v.y);
    p[4] = bytes01(v.z); p[5] = bytes23(v.z);
    p[6] = bytes01(v.w); p[7] = bytes23(v.w);
}
__device__ __forceinline__ uint4 pack16(const uint32_t (&p)[8])
{
    return make_uint4(pack4(p[0], p[1]), pack4(p[2], p[3]), pack4(p[4], p[5]), pack4(p[6], p[7]));
}

// DPP within a 16-lane row: lane i gets lane i-1 (shr) / i+1 (shl); 0 where
// the source is outside the row (one DPP move, no fill value to materialise)
__device__ __forceinline__ uint32_t row_shr1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
}
__device__ __forceinline__ uint32_t row_shl1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, true);
}
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_min(uint32_t v)
{
    return min(v, (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)v, CTRL, 0xF, 0xF, false));
}
// minimum over each aligned group of KL lanes (KL in 1,2,4,8,16), in every lane
template <int KL>
__device__ __forceinline__ uint32_t group_min(uint32_t v)
{
    if (KL >= 2) v = dpp_min<0xB1>(v);    // quad_perm [1,0,3,2]
    if (KL >= 4) v = dpp_min<0x4E>(v);    // quad_perm [2,3,0,1]
    if (KL >= 8) v = dpp_min<0x141>(v);   // row_half_mirror
    if (KL >= 16) v = dpp_min<0x140>(v);  // row_mirror
    return v;
}

// One chain is held by KL consecutive lanes; lane `lid` of the group owns
// chunks [lid*CH, lid*CH + CH) of the chain's Dp/16 chunks (16 disparities,
// 8 packed pairs each).  Chunks at or beyond nq are padding and stay guards.
template <int CH>
struct LaneState {
    uint32_t L[8 * CH];
    uint32_t mp;
};

// loads of one visited cell, issued a row ahead of their use
template <int CH, int NO>
struct RowBuf {
    uint4 m[CH];      // cost M
    uint4 p[CH];      // this family's forward (L - M), backward sweeps
    uint4 o[NO * CH]; // the other families' parts (final sweep)
};

struct ChainGeo {
    int lid;        // lane within the chain group
    int nq;         // valid chunks of the chain
    bool half_tail; // D % 16 == 8: the last valid chunk holds 4 real pairs
};

// one standard step over the lane's chunks.  m = cost pairs per chunk;
// returns (L - M) pairs in tp (ACC: adds them to tp).  trunc: 8-bit
// first-line truncation.
template <int CH, int KL, bool ACC, class MF, bool DEFER = false>
__device__ __forceinline__ void lane_step_f(LaneState<CH>& st, const ChainGeo& cg, MF&& mpair,
                                            uint32_t (&tp)[CH][8], bool trunc);

template <int CH, int KL, bool ACC = false>
__device__ __forceinline__ void lane_step(LaneState<CH>& st, const ChainGeo& cg, const uint32_t (&m)[CH][8],
                                          uint32_t (&tp)[CH][8], bool trunc)
{
    lane_step_f<CH, KL, ACC>(st, cg, [&](int q, int i) __attribute__((always_inline)) { return m[q][i]; }, tp, trunc);
}

// the same step with the cost pairs taken from packed bytes (4 per word, 16 per chunk).
// DEFER: the state is kept raw (minL' not subtracted): the new raw values are
// the normalized ones plus the previous raw minimum, which the caller tracks.
template <int CH, int KL, bool ACC = false, bool DEFER = false>
__device__ __forceinline__ void lane_step_pk(LaneState<CH>& st, const ChainGeo& cg, const uint4 (&mb)[CH],
                                             uint32_t (&tp)[CH][8])
{
    auto mf = [&](int q, int i) __attribute__((always_inline)) {
        const uint32_t w = i < 2 ? mb[q].x : i < 4 ? mb[q].y : i < 6 ? mb[q].z : mb[q].w;
        return (i & 1) ? bytes23(w) : bytes01(w);
    };
    lane_step_f<CH, KL, ACC, decltype(mf)&, DEFER>(st, cg, mf, tp, false);
}

template <int CH, int KL, bool ACC, class MF, bool DEFER>
__device__ __forceinline__ void lane_step_f(LaneState<CH>& st, const ChainGeo& cg, MF&& mpair,
                                            uint32_t (&tp)[CH][8], bool trunc)
{
    // left lane's last pair and right lane's first pair (old); KL divides the
    // 16-lane DPP row, so the lanes at a row's ends are group ends, which take
    // the guard instead
    const uint32_t nbl = row_shr1(st.L[8 * CH - 1]);
    const uint32_t nbr = row_shl1(st.L[0]);
    uint32_t prv = (cg.lid == 0) ? kGuard2 : nbl;
    const uint32_t rfill = (cg.lid == KL - 1) ? kGuard2 : nbr;
    const uint32_t mpp = dup16(st.mp);
    const us2 p2t = U2(mpp) + (us2){(unsigned short)kP2, (unsigned short)kP2};
    // truncation and padding as one and-or per pair: every new value is below
    // 0x4000 (at most P2 + 255 normalised, raw values stay below the guard), so
    // or-ing the guard 0x3FFF into a padding pair yields the guard itself
    const uint32_t tmask = trunc ? 0x00FF00FFu : 0xFFFFFFFFu;
    us2 rmin = U2(0xFFFFFFFFu);
    static_for<0, CH>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int qg = cg.lid * CH + q;
        const bool valid = qg < cg.nq;
        const bool ht = cg.half_tail && qg == cg.nq - 1;
        const uint32_t pad_lo = valid ? 0u : kGuard2;           // pairs 0..3 of the chunk
        const uint32_t pad_hi = valid && !ht ? 0u : kGuard2;    // pairs 4..7
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int j = 8 * q + i;
            const uint32_t cur = st.L[j];
            const uint32_t nxt = (j + 1 < 8 * CH) ? st.L[j + 1] : rfill;
            const us2 left = U2(__builtin_amdgcn_alignbit(cur, prv, 16));
            const us2 right = U2(__builtin_amdgcn_alignbit(nxt, cur, 16));
            const us2 nb = umin2(left, right) + (us2){(unsigned short)kP1, (unsigned short)kP1};
            const us2 t = umin2(umin2(nb, U2(cur)), p2t);
            const us2 d = DEFER ? t : t - U2(mpp);
            const us2 l = U2((W2(d + U2(mpair(q, i))) & tmask) | (i < 4 ? pad_lo : pad_hi));
            tp[q][i] = ACC ? W2(U2(tp[q][i]) + d) : W2(d);
            st.L[j] = W2(l);
            prv = cur;
            rmin = umin2(rmin, l);
        }
    });
    st.mp = group_min<KL>(min((uint32_t)rmin.x, (uint32_t)rmin.y));
}

// state <- given pairs (first-line c~ or a restart), padding as guards
template <int CH, int KL>
__device__ __forceinline__ void lane_set(LaneState<CH>& st, const ChainGeo& cg, const uint32_t (&m)[CH][8], bool on)
{
    us2 rmin = U2(0xFFFFFFFFu);
    static_for<0, CH>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int qg = cg.lid * CH + q;
        const bool valid = on && qg < cg.nq;
        const bool ht = cg.half_tail && qg == cg.nq - 1;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint32_t l = (!valid || (ht && i >= 4)) ? kGuard2 : m[q][i];
            st.L[8 * q + i] = l;
            rmin = umin2(rmin, U2(l));
        }
    });
    st.mp = group_min<KL>(min((uint32_t)rmin.x, (uint32_t)rmin.y));
    st.mp = st.mp == 0x3FFFu ? 0u : st.mp;   // a chain with no state yet: minL' = 0
}

template <int CH>
__device__ __forceinline__ void lane_guard(LaneState<CH>& st, bool keep)
{
    static_for<0, 8 * CH>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        st.L[j] = keep ? st.L[j] : kGuard2;
    });
    st.mp = keep ? st.mp : 0u;
}

__device__ __forceinline__ uint4 add_u16x8(uint4 a, uint4 b)
{
    return make_uint4(W2(U2(a.x) + U2(b.x)), W2(U2(a.y) + U2(b.y)), W2(U2(a.z) + U2(b.z)), W2(U2(a.w) + U2(b.w)));
}

// prefetch depth of the family kernels (cells whose loads are in flight per
// chain); overridable at build time for tuning.  At 16 lanes per chain the
// rows, columns and diagonals take 8 (a single rig has under one wave per
// SIMD, so the bytes in flight per wave set its bandwidth); the
// anti-diagonal family, whose cells also carry three other families' bytes,
// keeps 4 (8 would double its registers and halve its occupancy).
#ifndef SISTER_PF16
#define SISTER_PF16 8
#endif
#ifndef SISTER_PF16A
#define SISTER_PF16A 4
#endif
#ifndef SISTER_PF8
#define SISTER_PF8 2
#endif
#ifndef SISTER_PF4
#define SISTER_PF4 2
#endif
template <int CH, int KL, int FAM>
struct FamTune {
    static constexpr int PF = KL == 16 ? (FAM == 3 ? SISTER_PF16A : SISTER_PF16) : (KL == 8 ? SISTER_PF8 : SISTER_PF4);
};

// Family kernel: FAM 0 rows (lanes over rows, steps along x), 1 columns,
// 2 diagonals (k = x - y), 3 anti-diagonals (s = x + y, last, with the final
// winner); each chain swept forward (pass 0 direction) then backward.
template <int CH, int KL, int FAM>
__global__ __launch_bounds__(64) void k_fam(const RigDesc* __restrict__ rigs, int nrig, const int* __restrict__ bs,
                                            int mode, int keep_sum, int rsplit)
{
    // The rows family stores its two directions' (L - M) separately (forward
    // in part[0], backward in rowsb), so neither reads the other.  rsplit: the
    // two directions run as separate waves, the grid doubled and its second
    // half sweeping backward; a wave holding the top or bottom row's chain,
    // whose exact edge values add both directions, sweeps both in its
    // first-half block.  For the anti-diagonal family rsplit 1 / 2 runs only
    // the forward / backward (final) sweep: the forward one needs M alone, so
    // it can run beside the other families.
    extern __shared__ uint32_t lds_s[];   // final sweep: S pairs of the wave's chains
    constexpr int CPW = 64 / KL;          // chains per wave
    constexpr int NO = (FAM == 3) ? 4 : 1;
    int bid = blockIdx.x, dsel = -1;      // dsel: the one direction of this block, -1 both
    if (FAM == 0 && rsplit) {
        const int half = bs[nrig];
        dsel = bid >= half ? 1 : 0;
        bid -= dsel * half;
    }
    const int ri = find_rig(bs, nrig, bid);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x;
    const int grp = lane / KL;
    ChainGeo cg;
    cg.lid = lane % KL;
    cg.nq = g.Dp >> 4;
    cg.half_tail = (g.D & 15) != 0;
    const int Wp = g.Wp, Hp = g.Hp, Dp = g.Dp, D = g.D;
    const int c = (bid - bs[ri]) * CPW + grp;     // chain index
    const int nchain = FAM == 0 ? Hp : (FAM == 1 ? Wp : Wp + Hp - 1);
    const bool valid = c < nchain;
    const int cw0 = (bid - bs[ri]) * CPW;
    const int cw1 = min(cw0 + CPW - 1, nchain - 1);
    uint8_t* part = g.part[FAM];
    uint16_t* edge = g.edge + (size_t)FAM * 2 * (size_t)Wp * Dp;
    const int boff = cg.lid * CH * 16;     // byte offset of the lane's chunks in a cell
    const bool final_pass = FAM == 3;
    // rows family: byte offset of the lane's chunks in column 0 of its row
    const size_t row_off = FAM == 0 ? (size_t)(valid ? c : 0) * Wp * Dp + boff : 0;
    // other families: the chain's cell at row y lies at y (W' + s) Dp + lane_off,
    // the row term uniform over the wave (s = 0 columns, +1 diagonals, -1
    // anti-diagonals; lane_off may be negative where the chain starts left of
    // column 0, and is used only at active cells)
    const long ystride = (long)(Wp + (FAM == 2 ? 1 : FAM == 3 ? -1 : 0)) * Dp;
    const long lane_off = (long)(FAM == 2 ? c - (Hp - 1) : c) * Dp + boff;

    // ---- chain geometry --------------------------------------------------
    // steps 0..n-1 of the wave; cell of chain c at step t, per direction
    int nsteps;
    int ylo = 0, yhi = 0, ent_dn = 0, ent_up = 0;
    if (FAM == 0) {
        nsteps = Wp;
    } else if (FAM == 1) {
        ylo = 0; yhi = Hp - 1; ent_dn = 0; ent_up = Hp - 1;
        nsteps = Hp;
    } else if (FAM == 2) {
        const int klo = cw0 - (Hp - 1), khi = cw1 - (Hp - 1);
        ylo = max(0, -khi); yhi = min(Hp - 1, Wp - 1 - klo);
        ent_dn = max(0, -klo); ent_up = min(Hp - 1, Wp - 1 - khi);
        nsteps = yhi - ylo + 1;
    } else {
        ylo = max(0, cw0 - (Wp - 1)); yhi = min(Hp - 1, cw1);
        ent_dn = max(0, cw1 - (Wp - 1)); ent_up = min(Hp - 1, cw0);
        nsteps = yhi - ylo + 1;
    }
    auto cell_of = [&](int dir, int t, int& y, int& x) {
        if (FAM == 0) { y = c; x = dir == 0 ? t : Wp - 1 - t; return; }
        y = dir == 0 ? ylo + t : yhi - t;
        if (FAM == 1) x = c;
        else if (FAM == 2) x = y + (c - (Hp - 1));
        else x = c - y;
    };
    auto is_active = [&](int y, int x) { return valid && x >= 0 && x < Wp && y >= 0 && y < Hp; };

    const bool top = FAM == 0 && c == 0, bot = FAM == 0 && c == Hp - 1;
    const bool want_sub = final_pass && mode == 0 && g.subpix != nullptr;
    uint16_t* outm = g.out[mode];

    int dlo = 0, dhi = 1;
    if (FAM == 3 && rsplit) dlo = dhi = rsplit - 1;
    if (dsel >= 0) {
        const bool edge_wave = cw0 == 0 || cw1 >= Hp - 1;
        if (edge_wave) {
            if (dsel == 1) return;
        } else {
            dlo = dhi = dsel;
        }
    }
    LaneState<CH> st;
    for (int dir = dlo; dir <= dhi; dir++) {
        lane_guard<CH>(st, false);
        const bool need_p = dir == 1 && FAM != 0;
        const bool need_o = final_pass && dir == 1;
        auto fetch = [&](int t, RowBuf<CH, NO>& b) __attribute__((always_inline)) {
            if constexpr (FAM == 0) {
                // rows: the lane's row offset is fixed, the column uniform
                const int x = dir == 0 ? t : Wp - 1 - t;
                const uint8_t* src = g.msum + row_off + (size_t)x * Dp;
                static_for<0, CH>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const bool ok = t < nsteps && valid && cg.lid * CH + q < cg.nq;
                    // not loaded: a padding chunk (its pairs take the guard whatever
                    // bytes they hold), a chain past the rig (never stored) or a
                    // step past the row (never processed): no zero fill
                    if (ok) b.m[q] = LD16(src + 16 * q);
                });
                return;
            }
            int y, x;
            cell_of(dir, t, y, x);
            const bool act = t < nsteps && is_active(y, x);
            const size_t off = (size_t)((long)y * ystride + lane_off);
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const bool ok = act && cg.lid * CH + q < cg.nq;
                b.m[q] = ok ? LD16(g.msum + off + 16 * q) : make_uint4(0, 0, 0, 0);
                if (need_p) b.p[q] = ok ? LD16(part + off + 16 * q) : make_uint4(0, 0, 0, 0);
                if (need_o) {
                    b.o[q] = ok ? LD16(g.part[0] + off + 16 * q) : make_uint4(0, 0, 0, 0);
                    if constexpr (NO == 4) {
                        b.o[CH + q] = ok ? LD16(g.part[1] + off + 16 * q) : make_uint4(0, 0, 0, 0);
                        b.o[2 * CH + q] = ok ? LD16(g.part[2] + off + 16 * q) : make_uint4(0, 0, 0, 0);
                        b.o[3 * CH + q] = ok ? LD16(g.rowsb + off + 16 * q) : make_uint4(0, 0, 0, 0);
                    }
                }
            });
        };
        auto process = [&](int t, RowBuf<CH, NO>& b) __attribute__((always_inline)) {
            int y, x;
            cell_of(dir, t, y, x);
            const bool act = is_active(y, x);
            const size_t off = (size_t)((long)y * ystride + lane_off);   // used at active cells
            uint32_t m[CH][8], tp[CH][8];
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                unpack16(b.m[q], m[q]);
            });
            // ---------------- rows family ----------------
            if (FAM == 0) {
                const bool first_line = dir == 0 ? top : bot;   // 255 -> 0, 8-bit truncation
                if (first_line) {
                    static_for<0, CH>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
#pragma unroll
                        for (int i = 0; i < 8; i++) m[q][i] = tilde2(m[q][i]);
                    });
                }
                if (t == 0) {
                    lane_set<CH, KL>(st, cg, m, true);      // L = M (c~ on a first line), contributes
                    static_for<0, CH>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
#pragma unroll
                        for (int i = 0; i < 8; i++) tp[q][i] = 0;
                    });
                } else {
                    lane_step<CH, KL>(st, cg, m, tp, first_line);
                }
                if (!act) return;
                const bool edge_row = top || bot;
                // its bytes are read only where S is: at the crop, or everywhere
                // with the debug sums
                const bool crop = keep_sum || (y >= D && y < D + g.H && x >= D && x < D + g.W);
                static_for<0, CH>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if (cg.lid * CH + q >= cg.nq) return;
                    if (edge_row) {
                        uint16_t* e = edge + ((size_t)(top ? 0 : 1) * Wp + x) * Dp + (cg.lid * CH + q) * 16;
                        uint4 a = make_uint4(st.L[8 * q + 0], st.L[8 * q + 1], st.L[8 * q + 2], st.L[8 * q + 3]);
                        uint4 bb = make_uint4(st.L[8 * q + 4], st.L[8 * q + 5], st.L[8 * q + 6], st.L[8 * q + 7]);
                        if (cg.half_tail && cg.lid * CH + q == cg.nq - 1) bb = make_uint4(0, 0, 0, 0);
                        if (dir == 1) {
                            a = add_u16x8(a, LD16(e));
                            bb = add_u16x8(bb, LD16(e + 8));
                        }
                        ST16(e, a);
                        ST16(e + 8, bb);
                    } else if (crop) {
                        ST16((dir == 1 ? g.rowsb : part) + row_off + (size_t)x * Dp + 16 * q, pack16(tp[q]));
                    }
                });
                return;
            }
            // ---------------- vertical families ----------------
            const int first_row = dir == 0 ? 0 : Hp - 1;
            if (y == first_row) {
                // first line of the pass: state c~, no contribution (sgm.cpp:103-207)
                static_for<0, CH>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
#pragma unroll
                    for (int i = 0; i < 8; i++) m[q][i] = tilde2(m[q][i]);
                });
                lane_set<CH, KL>(st, cg, m, act);
                if (final_pass && keep_sum && g.sum && act) {
                    // S of row H'-1 = the four families' exact edge rows
                    static_for<0, CH>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if (cg.lid * CH + q >= cg.nq) return;
                        const size_t eo = (size_t)x * Dp + cg.lid * CH * 16 + 16 * q;
#pragma unroll
                        for (int h = 0; h < 2; h++) {
                            uint4 acc = make_uint4(0, 0, 0, 0);
#pragma unroll
                            for (int fm = 0; fm < 4; fm++)
                                acc = add_u16x8(acc, LD16(g.edge + (size_t)(fm * 2 + 1) * Wp * Dp + eo + 8 * h));
                            ST16(g.sum + ((size_t)y * Wp) * Dp + eo + 8 * h, acc);
                        }
                    });
                }
                return;
            }
            lane_step<CH, KL>(st, cg, m, tp, false);
            const bool entering = dir == 0 ? (y <= ent_dn) : (y >= ent_up);
            if (entering) lane_guard<CH>(st, act);
            const bool edge_row = (y == 0 || y == Hp - 1);
            const int oy = y - D, ox = x - D;
            const bool crop = final_pass && act && oy >= 0 && oy < g.H && ox >= 0 && ox < g.W;
            uint32_t key = 0xFFFFFFFFu;
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int qg = cg.lid * CH + q;
                if (!act || qg >= cg.nq) return;
                if (edge_row) {
                    uint16_t* e = edge + ((size_t)(y == 0 ? 0 : 1) * Wp + x) * Dp + qg * 16;
                    uint4 a = make_uint4(st.L[8 * q + 0], st.L[8 * q + 1], st.L[8 * q + 2], st.L[8 * q + 3]);
                    uint4 bb = make_uint4(st.L[8 * q + 4], st.L[8 * q + 5], st.L[8 * q + 6], st.L[8 * q + 7]);
                    if (cg.half_tail && qg == cg.nq - 1) bb = make_uint4(0, 0, 0, 0);
                    ST16(e, a);
                    ST16(e + 8, bb);
                    if (final_pass && keep_sum && g.sum && y == 0) {
                        // S of row 0: the other families' edge rows + this sweep
                        const size_t eo = (size_t)x * Dp + qg * 16;
                        uint4 s0 = a, s1 = bb;
#pragma unroll
                        for (int fm = 0; fm < 3; fm++) {
                            s0 = add_u16x8(s0, LD16(g.edge + (size_t)(fm * 2) * Wp * Dp + eo));
                            s1 = add_u16x8(s1, LD16(g.edge + (size_t)(fm * 2) * Wp * Dp + eo + 8));
                        }
                        ST16(g.sum + eo, s0);
                        ST16(g.sum + eo + 8, s1);
                    }
                    return;
                }
                if (dir == 0) {
                    ST16(part + off + 16 * q, pack16(tp[q]));
                    return;
                }
                uint32_t pf[8];
                unpack16(b.p[q], pf);
#pragma unroll
                for (int i = 0; i < 8; i++) tp[q][i] = W2(U2(tp[q][i]) + U2(pf[i]));
                if (!final_pass) {
                    ST16(part + off + 16 * q, pack16(tp[q]));
                    return;
                }
                if (!(crop || keep_sum)) return;
                // S = 8 M + part[rows] + part[cols] + part[diag] + this family
                uint32_t pr[8], pc[8], pd[8], S[8];
                const uint4 rf = b.o[q], rb = b.o[NO == 4 ? 3 * CH + q : q];
                unpack16(make_uint4(rf.x + rb.x, rf.y + rb.y, rf.z + rb.z, rf.w + rb.w), pr);   // bytes <= 200
                unpack16(b.o[NO == 4 ? CH + q : q], pc);
                unpack16(b.o[NO == 4 ? 2 * CH + q : q], pd);
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const us2 s8 = U2(m[q][i]) << (us2){3, 3};
                    S[i] = W2(s8 + U2(pr[i]) + U2(pc[i]) + U2(pd[i]) + U2(tp[q][i]));
                }
                if (keep_sum && g.sum) {
                    uint16_t* sc = g.sum + (off - boff) + qg * 16;
                    ST16(sc, make_uint4(S[0], S[1], S[2], S[3]));
                    ST16(sc + 8, make_uint4(S[4], S[5], S[6], S[7]));
                }
                if (!crop) return;
#pragma unroll
                for (int i = 0; i < 8; i++) {
                    const uint32_t d0 = (uint32_t)(qg * 16 + 2 * i);
                    const uint32_t k0 = ((S[i] & 0xFFFFu) << 9) | d0;
                    const uint32_t k1 = ((S[i] >> 16) << 9) | (d0 + 1);
                    key = min(key, d0 < (uint32_t)D ? k0 : 0xFFFFFFFFu);
                    key = min(key, d0 + 1 < (uint32_t)D ? k1 : 0xFFFFFFFFu);
                    if (want_sub) lds_s[(grp * cg.nq * 8) + qg * 8 + i] = S[i];
                }
            });
            if (final_pass && dir == 1) {
                // the winner of each chain's cell: minimum key over the group
                key = group_min<KL>(key);
                if (want_sub) __syncthreads();
                if (crop && cg.lid == 0) {
                    const uint32_t best = key & 511u;
                    const size_t o = (size_t)oy * g.W + ox;
                    if (outm) G(outm)[o] = (uint16_t)min(best * 255u, 65535u);
                    if (mode == 0 && g.disp) G(g.disp)[o] = (uint16_t)best;
                    if (want_sub) {
                        float delta = 0.0f;
                        if (best > 0 && (int)best < D - 1) {
                            const uint16_t* sv = (const uint16_t*)(lds_s + grp * cg.nq * 8);
                            const int sm = sv[best - 1], s0 = (int)(key >> 9), sp = sv[best + 1];
                            const int den = sm - 2 * s0 + sp;
                            if (den > 0) delta = (float)(sm - sp) / (2.0f * (float)den);
                        }
                        G(g.subpix)[o] = (float)best + delta;
                    }
                }
                if (want_sub) __syncthreads();
            }
        };
        // software pipeline: a ring of PF cell buffers, the loads of the next
        // PF - 1 cells in flight while the current one is computed
        constexpr int PF = FamTune<CH, KL, FAM>::PF;
        RowBuf<CH, NO> ring[PF];
        static_for<0, PF - 1>([&](auto ic) __attribute__((always_inline)) {
            fetch(decltype(ic)::value, ring[decltype(ic)::value]);
        });
        for (int t = 0; t < nsteps; t += PF) {
            static_for<0, PF>([&](auto ic) __attribute__((always_inline)) {
                constexpr int i = decltype(ic)::value;
                fetch(t + i + PF - 1, ring[(i + PF - 1) % PF]);
                if (t + i < nsteps) process(t + i, ring[i]);
            });
        }
    }
}

// ---------------------------------------------------------------------------
// K6b fused vertical families: the column, diagonal and anti-diagonal chains
// of a pass in one sweep (same recurrence and pass rules as k_fam).  Lanes own
// columns (kVKL = 4 lanes per column, CH chunks of 16 disparities each), so
// every cell of a row sees its three vertical paths in one lane group:
//   downward pass:  P = rows_fwd + rows_bwd + sum_f (L_f,down - M)  (u16, rows 1..H'-2)
//   upward pass:    S = 8 M + P + sum_f (L_f,up - M), winner and sub-pixel
// -- 8 bytes per cell over both passes, against 17 for the three separate
// families (5 + 5 + 7).  Along a pass the diagonal-family state moves one
// column per row (down-right, up-left) and the anti-diagonal state the other
// way; both live in LDS, one slot per column, updated in place between
// barriers.  Call FL the family whose state moves to higher columns (diagonal
// going down, anti-diagonal going up) and FR the other.
// A block owns ncol columns and also computes kVGhost ghost columns on each
// side (one wave each; what matters there is the family that flows inwards,
// FL on the left and FR on the right).  A wrong state entering at the ghost
// edge moves inwards one column per row, so the owned columns stay exact for
// kVGhost rows without any exchange with neighbouring blocks: one launch runs
// one segment of kVGhost rows, and the states cross segments through memory
// (owned columns stored at the end, owned + ghost columns loaded at the start).
// No block ever waits for another.
// Buffers: the rows family's two directions in part[0] and rowsb (bytes,
// added as words: each sum is at most 200), P (u16) in part[1..2], or over
// the rows bytes once the down pass has read them in plans short of memory
// (pd_at below), the segment states in vstate
// ([parity][cols, diag, anti][W'][Dp + 16] bytes: the normalised (L - M)
// bytes, then minL').
// ---------------------------------------------------------------------------
#ifndef SISTER_VS_XCD
#define SISTER_VS_XCD 1
#endif
constexpr int kVGhost = 16;        // ghost columns per side = rows per launch
constexpr int kVMaxThreads = 768;  // 12 waves: up to 168 VGPRs per lane
// Lanes per column: 4 up to D = 192 (3 chunks of 16 per lane), 8 up to D =
// 384.  Owned columns per block: 144 with 4 lanes (176 computed, 11 waves;
// 16 blocks per 2304-wide rig), 64 with 8 (96 computed, 12 waves).
__host__ __device__ constexpr int vs_cols(int KL) { return KL == 4 ? 144 : 64; }
// 8-byte granules per pair slot of the family state slots: nt computed
// columns plus the two border columns, KL lanes each, rounded up to 64
__host__ __device__ constexpr int vs_gs(int nt, int KL) { return ((nt + 2) * KL + 63) / 64 * 64; }

template <int CH>
__device__ __forceinline__ void vs_guard_all(LaneState<CH>& st)
{
    static_for<0, 8 * CH>([&](auto jc) { st.L[decltype(jc)::value] = kGuard2; });
    st.mp = 0;
}

// S pair j (0 <= j < 8 CH) of the lane, by a select tree
template <int CH>
__device__ __forceinline__ uint32_t vs_pick(const uint32_t (&S)[CH][8], int j)
{
    uint32_t r = 0;
    static_for<0, CH>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int i = j - 8 * q;
        const uint32_t a0 = (i & 1) ? S[q][1] : S[q][0], a1 = (i & 1) ? S[q][3] : S[q][2];
        const uint32_t a2 = (i & 1) ? S[q][5] : S[q][4], a3 = (i & 1) ? S[q][7] : S[q][6];
        const uint32_t b0 = (i & 2) ? a1 : a0, b1 = (i & 2) ? a3 : a2;
        const uint32_t v = (i & 4) ? b1 : b0;
        r = (i >= 0 && i < 8) ? v : r;
    });
    return r;
}

template <int KL>
__device__ __forceinline__ uint32_t group_max(uint32_t v)
{
    return ~group_min<KL>(~v);
}

// winner key (S << 9) | D of the low (HI = false) or high u16 half S of a pair,
// in one v_mad_u32_u16 (op_sel picks the half; D <= 64 is an inline constant,
// 512 comes in a register: one scalar operand at most on this VOP3)
template <int D, bool HI>
__device__ __forceinline__ uint32_t vs_key(uint32_t pair, uint32_t m512)
{
    static_assert(D >= 0 && D <= 64, "inline constant");
    uint32_t r;
    if (HI) asm("v_mad_u32_u16 %0, %1, %2, %3 op_sel:[1,0,0,0]" : "=v"(r) : "v"(pair), "v"(m512), "i"(D));
    else asm("v_mad_u32_u16 %0, %1, %2, %3" : "=v"(r) : "v"(pair), "v"(m512), "i"(D));
    return r;
}

template <int CH, int DIR, bool KEEP, bool FULL, int VKL>
__global__ __launch_bounds__(kVMaxThreads) void k_vsweep(const RigDesc* __restrict__ rigs, int nrig,
                                                        const int* __restrict__ bs, int ncol, int seg, int mode)
{
    // KEEP: also the debug sums S (all rows) and the edge rows they need.
    // FULL: D == Dp == 16 VKL CH, every chunk of every lane is real (no padding
    // chunk, no half chunk), so no per-chunk lane conditions are needed.
    constexpr bool keep_sum = KEEP;
    // deferred normalisation: within a segment the states stay raw (the step
    // does not subtract minL'); the row's output subtracts the three
    // families' pre-step minima once, and the segment-end store normalises.
    // Raw values grow by at most 255 per row, so 16 rows stay far below the
    // guard 0x3FFF.  The debug variant keeps exact states for its edge rows.
    constexpr bool DEFER = !KEEP;
    constexpr int KL = VKL, NP = 8 * CH;
    constexpr int VC = vs_cols(KL);                // owned columns per block
    extern __shared__ uint32_t xs[];               // FL slots, then FR slots
    // neighbouring column blocks on one XCD: a block's ghost columns are its
    // neighbours' owned ones, whose cost bytes and states then come from the
    // same L2
#if SISTER_VS_XCD
    const int bid = xcd_order(blockIdx.x, gridDim.x);
#else
    const int bid = blockIdx.x;
#endif
    const int ri = find_rig(bs, nrig, bid);
    const RigDesc& g = rigs[ri];
    const int Wp = g.Wp, Hp = g.Hp, D = g.D, Dp = g.Dp;
    const int t0 = seg * kVGhost;
    if (t0 >= Hp) return;                          // the rig has no such segment: whole block
    const int nrows = min(kVGhost, Hp - t0);
    (void)ncol;                                    // VC (compile-time LDS strides)
    constexpr int nt = VC + 2 * kVGhost;
    const int tid = threadIdx.x, c = tid / KL;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), nwaves = (int)(blockDim.x >> 6);
    const int x = (bid - bs[ri]) * VC - kVGhost + c;
    const bool act = x >= 0 && x < Wp;
    // every column of the wave inside the image: no border guards to apply
    const bool wave_act = __builtin_amdgcn_ballot_w64(!act) == 0;
    // ghost waves step all three families too (their outward-flowing chains
    // and columns are never read by an owned column), so no wave branches
    // around a step; the busiest SIMD still carries three owned waves.  The
    // kVGhost ghost columns of a side fill kVGhost KL / 64 waves.
    constexpr int GW = kVGhost * KL / 64;
    static_assert(GW * 64 == kVGhost * KL, "ghost columns fill whole waves");
    const bool own = wave >= GW && wave < nwaves - GW;
    ChainGeo cg;
    cg.lid = tid % KL;
    cg.nq = FULL ? KL * CH : Dp >> 4;
    cg.half_tail = FULL ? false : (D & 15) != 0;
    const int boff = cg.lid * CH * 16;
    // LDS slots of a family, lane-interleaved so that a wave's accesses are
    // contiguous: pairs 2m, 2m+1 of lane (column cc, lid) in the 8-byte granule
    // m * GS + (cc + 1) * KL + lid, then minL' of column cc at word
    // NP * GS + cc + 1.  Columns -1 and nt hold a border state (guards,
    // minL' = 0), written once and never overwritten, so reads need no test.
    // GS is a multiple of 64 granules, so that the compiler pairs the slot
    // accesses (ds_read2st64 / ds_write2st64).
    constexpr int GS = vs_gs(nt, KL);
    uint32_t* XL = xs;
    uint32_t* XR = xs + NP * GS + nt + 2;
    constexpr int famL = DIR == 0 ? 2 : 3, famR = DIR == 0 ? 3 : 2;   // family ids of FL, FR
    // segment states as bytes: at the segment's last row r a chain's
    // t = L - M(r) lies in [m, m + P2] (m = its row r-1 minimum, reached at that
    // row's argmin), so t - min_d t is the normalised (L - M) in [0, 100].  A
    // column's state is those Dp bytes and the u32 min_d L - min_d t, SB bytes
    // in all; the next segment rebuilds the normalised L = M(r) + byte.
    const int SB = Dp + 16;
    uint8_t* const Sg = (uint8_t*)g.vstate;
    const uint8_t* const msum = g.msum;
    auto sg_col = [&](int par, int f) { return Sg + ((size_t)(par * 3 + f - 1) * Wp + x) * SB; };
    // FULL: loads of a column outside the image read the nearest inside one
    // (in bounds, no exec-masked loads or zero fills in the row loop); such a
    // column's states are reset to guards and nothing of it is stored
    const int xc = min(max(x, 0), Wp - 1);
    // P is this kernel's own intermediate (rows 1..H'-2).  When the plan gives
    // it volumes of its own (part[1..2]): FULL, planes per (row, chunk slot
    // q, half h) in which lane (x, lid) holds 8 u16 at byte (x KL + lid) * 16,
    // so a wave's 64 lanes move 1 KB contiguous per access (3 ms faster than
    // the cell layout at 16 rigs); otherwise cell layout (chunk qg of cell
    // (y, x) at (y W' + x) Dp + 16 qg).
    // Plans short of memory give it none (part[1] null, FULL only): P then
    // takes the place of the rows family's bytes of the same row and columns,
    // which the down pass reads just before it stores P.  The 2 Dp bytes of
    // column x occupy its Dp part[0] bytes and its Dp rowsb bytes, and only
    // the wave that owns column x reads or writes them; its loads of the row's
    // bytes complete before its P stores (P depends on them).  A lane's
    // 16-byte piece k = 2 q + h lies in part[0] for k < CH, in rowsb above.
    // Waves whose columns all lie inside the row keep whole-wave planes:
    // plane k of the wave's 64 / KL columns is 1 KB with wave lane l at byte
    // 16 l, planes 0..CH-1 over the wave's part[0] bytes of the row and
    // CH..2CH-1 over its rowsb bytes.  The others (the last wave of a row)
    // keep a column's pieces together: piece k of lane lid at 16 (lid CH +
    // k mod CH) in the column's bytes of its half.  (1.3 ms slower per
    // 64-rig step than volumes of its own.)  Either way the address is a
    // uniform base plus a loop-invariant lane offset.
    uint16_t* const Pd = (uint16_t*)g.part[1];
    const bool p_own = Pd != nullptr;
    constexpr int CPW = 64 / KL;                       // columns per wave
    const int lw = tid & 63;
    const int x0w = __builtin_amdgcn_readfirstlane(x - lw / KL);   // the wave's first column
    // S, and so P and the rows family's bytes it is formed from, is read only
    // at the crop (D <= y < D + H, D <= x < D + W) outside the debug variant:
    // rows outside it, and waves whose columns all lie outside it, neither
    // load nor store them (the family steps run everywhere)
    const bool wave_crop = KEEP || (x0w + CPW > D && x0w < D + g.W);
    const bool wave_planes = FULL && x0w >= 0 && x0w + CPW <= Wp;
    const uint32_t pd_lane = p_own ? (uint32_t)(xc * KL + cg.lid) * 16u
                           : wave_planes ? (uint32_t)lw * 16u : (uint32_t)(xc * Dp + cg.lid * CH * 16);
    auto pd_at = [&](int y, int q, int h) __attribute__((always_inline)) {
        if (FULL) {
            if (p_own) {
                const uint16_t* plane = Pd + (((size_t)y * CH + q) * 2 + h) * ((size_t)Wp * KL * 8);   // uniform
                return (uint16_t*)((const char*)plane + pd_lane);
            }
            const int k = 2 * q + h, r = k >= CH ? 1 : 0, kk = k - r * CH;
            const size_t s = (size_t)y * Wp * Dp + (wave_planes ? (size_t)x0w * Dp + kk * 1024 : (size_t)kk * 16);
            return (uint16_t*)((r ? g.rowsb : g.part[0]) + s + pd_lane);
        }
        return Pd + ((size_t)y * Wp + x) * Dp + boff + 16 * q + 8 * h;
    };

    // lane state <-> a column's segment state (bytes, then minL' at byte Dp)
    // mr: the cost bytes M(r) of the segment's last row r (the previous
    // segment's, when loading); padding pairs come back as guards
    auto st_load = [&](LaneState<CH>& st, const uint8_t* v, const uint4 (&mr)[CH]) __attribute__((always_inline)) {
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int qg = cg.lid * CH + q;
            const bool ok = act && (FULL || qg < cg.nq);
            const bool tail = !FULL && cg.half_tail && qg == cg.nq - 1;
            uint32_t m[8], bt[8];
            unpack16(mr[q], m);
            unpack16(ok ? LD16(v + qg * 16) : make_uint4(0, 0, 0, 0), bt);
#pragma unroll
            for (int k = 0; k < 8; k++)
                st.L[8 * q + k] = (!ok || (tail && k >= 4)) ? kGuard2 : W2(U2(m[k]) + U2(bt[k]));
        });
        st.mp = act ? G((const uint32_t*)(v + Dp))[0] : 0u;
    };
    auto st_store = [&](const LaneState<CH>& st, uint8_t* v, const uint4 (&mr)[CH]) __attribute__((always_inline)) {
        uint32_t tq[CH][8];
        us2 tm = U2(0xFFFFFFFFu);
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            uint32_t m[8];
            unpack16(mr[q], m);
#pragma unroll
            for (int k = 0; k < 8; k++) {
                tq[q][k] = W2(U2(st.L[8 * q + k]) - U2(m[k]));   // padding pairs: guard-sized
                tm = umin2(tm, U2(tq[q][k]));
            }
        });
        const uint32_t tmin = group_min<KL>(min((uint32_t)tm.x, (uint32_t)tm.y));
        const us2 n = U2(dup16(tmin));
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int qg = cg.lid * CH + q;
            if (!(FULL || qg < cg.nq)) return;
            uint32_t b[8];
#pragma unroll
            for (int k = 0; k < 8; k++) b[k] = W2(U2(tq[q][k]) - n);
            ST16(v + qg * 16, pack16(b));
        });
        if (cg.lid == 0) G((uint32_t*)(v + Dp))[0] = st.mp - tmin;
    };
    auto slot_write = [&](uint32_t* X, const LaneState<CH>& st) __attribute__((always_inline)) {
        uint2* w = (uint2*)X + (c + 1) * KL + cg.lid;
        static_for<0, NP / 2>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            w[k * GS] = make_uint2(st.L[2 * k], st.L[2 * k + 1]);
        });
        if (cg.lid == 0) X[NP * GS + c + 1] = st.mp;
    };
    auto slot_read = [&](const uint32_t* X, int cc, LaneState<CH>& st) __attribute__((always_inline)) {
        const uint2* w = (const uint2*)X + (cc + 1) * KL + cg.lid;   // -1 <= cc <= nt
        static_for<0, NP / 2>([&](auto kc) {
            constexpr int k = decltype(kc)::value;
            const uint2 v = w[k * GS];
            st.L[2 * k] = v.x;
            st.L[2 * k + 1] = v.y;
        });
        st.mp = X[NP * GS + cc + 1];
    };
    auto row_of = [&](int t) { return DIR == 0 ? t : Hp - 1 - t; };
    auto load_m = [&](int t, uint4 (&b)[CH]) __attribute__((always_inline)) {
        const size_t cell = ((size_t)row_of(t) * Wp + (FULL ? xc : x)) * Dp + boff;
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool ok = act && (FULL || cg.lid * CH + q < cg.nq);
            b[q] = FULL ? LD16(msum + cell + 16 * q) : ok ? LD16(msum + cell + 16 * q) : make_uint4(0, 0, 0, 0);
        });
    };
    // exact L of a family on an edge row (debug sums only), as k_fam stores it
    auto edge_out = [&](int f, int y, const LaneState<CH>& st) __attribute__((always_inline)) {
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int qg = cg.lid * CH + q;
            if (qg >= cg.nq) return;
            uint16_t* e = g.edge + ((size_t)(f * 2 + (y == 0 ? 0 : 1)) * Wp + x) * Dp + qg * 16;
            const uint4 a = make_uint4(st.L[8 * q + 0], st.L[8 * q + 1], st.L[8 * q + 2], st.L[8 * q + 3]);
            uint4 bb = make_uint4(st.L[8 * q + 4], st.L[8 * q + 5], st.L[8 * q + 6], st.L[8 * q + 7]);
            if (cg.half_tail && qg == cg.nq - 1) bb = make_uint4(0, 0, 0, 0);
            ST16(e, a);
            ST16(e + 8, bb);
        });
    };
    // S of an edge row (debug): the four families' exact rows summed
    auto edge_sum = [&](int y) __attribute__((always_inline)) {
        const int e = y == 0 ? 0 : 1;
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int qg = cg.lid * CH + q;
            if (qg >= cg.nq) return;
            const size_t eo = (size_t)x * Dp + qg * 16;
#pragma unroll
            for (int h = 0; h < 2; h++) {
                uint4 acc = make_uint4(0, 0, 0, 0);
#pragma unroll
                for (int fm = 0; fm < 4; fm++)
                    acc = add_u16x8(acc, LD16(g.edge + (size_t)(fm * 2 + e) * Wp * Dp + eo + 8 * h));
                ST16(g.sum + (size_t)y * Wp * Dp + eo + 8 * h, acc);
            }
        });
    };

    const bool want_sub = DIR == 1 && mode == 0 && g.subpix != nullptr;
    uint16_t* const outm = g.out[mode];
    LaneState<CH> sc, sx;
    uint4 mb[CH];
    // the border slots (columns -1 and nt) of both families
    for (int i = tid; i < 2 * (NP / 2) * KL; i += blockDim.x) {
        const int side = i / ((NP / 2) * KL), rem = i % ((NP / 2) * KL);
        const int pos = side ? nt + 1 : 0, k = rem / KL, lid = rem % KL;
        ((uint2*)XL)[k * GS + pos * KL + lid] = make_uint2(kGuard2, kGuard2);
        ((uint2*)XR)[k * GS + pos * KL + lid] = make_uint2(kGuard2, kGuard2);
        if (k == 0 && lid == 0) {
            XL[NP * GS + pos] = 0;
            XR[NP * GS + pos] = 0;
        }
    }
    if (t0 > 0) {
        const int pin = (seg - 1) & 1;
        uint4 mr[CH];
        load_m(t0 - 1, mr);
        st_load(sc, sg_col(pin, 1), mr);
        st_load(sx, sg_col(pin, famL), mr);
        slot_write(XL, sx);
        st_load(sx, sg_col(pin, famR), mr);
        slot_write(XR, sx);
    } else {
        vs_guard_all<CH>(sc);
    }
    load_m(t0, mb);
    __syncthreads();
#pragma unroll 1
    for (int i = 0; i < nrows; i++) {
        const int t = t0 + i, y = row_of(t);
        uint32_t acc[CH][8];
        uint4 mc[CH];
        static_for<0, CH>([&](auto qc) { mc[decltype(qc)::value] = mb[decltype(qc)::value]; });
        if (i + 1 < nrows) load_m(t + 1, mb);
        if (t == 0) {
            // first line of the pass: every chain takes c~ and adds nothing (sgm.cpp:103-207)
            uint32_t m[CH][8];
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                unpack16(mc[q], m[q]);
#pragma unroll
                for (int k = 0; k < 8; k++) m[q][k] = tilde2(m[q][k]);
            });
            lane_set<CH, KL>(sc, cg, m, act);
            lane_set<CH, KL>(sx, cg, m, act);
            slot_write(XL, sx);
            slot_write(XR, sx);
            if (DIR == 1 && keep_sum && g.sum && own && act) edge_sum(y);
            __syncthreads();
            continue;
        }
        const bool interior = y >= 1 && y <= Hp - 2;
        const bool edge_row = !interior;
        const bool pc = wave_crop && (KEEP || (y >= D && y < D + g.H));   // P and rows bytes needed
        const size_t cell = ((size_t)y * Wp + x) * Dp + boff;
        // addends of the row's output, loaded now and used after the three steps
        uint4 ob[2 * CH];
        static_for<0, 2 * CH>([&](auto kc) { ob[decltype(kc)::value] = make_uint4(0, 0, 0, 0); });
        if (own && act && interior && pc) {
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if (!FULL && cg.lid * CH + q >= cg.nq) return;
                if (DIR == 0) {
                    ob[q] = LD16(g.part[0] + cell + 16 * q);
                    ob[CH + q] = LD16(g.rowsb + cell + 16 * q);
                } else {
                    ob[2 * q] = LD16(pd_at(y, q, 0));
                    ob[2 * q + 1] = LD16(pd_at(y, q, 1));
                }
            });
        }
        // (the steps' L - M outputs are not kept: their sum is formed after the
        // three steps from the new states, sum_f (L_f - M) = sum_f L_f - 3 M,
        // which keeps 24 registers free while the states are stepped)
        // Two barriers per row: FL slots are read (old) before the first and
        // written after it; FR slots read between the two and written after
        // the second.  The next row's FL reads follow this row's second
        // barrier, its FR reads its own first one, which orders them after
        // this row's writes.
        uint32_t dead[CH][8];
        slot_read(XL, c - 1, sx);
        __syncthreads();
        uint32_t off;   // DEFER: sum of the three pre-step minima
        {
            off = sx.mp;
            lane_step_pk<CH, KL, false, DEFER>(sx, cg, mc, dead);
            if (!wave_act) lane_guard<CH>(sx, act);
            slot_write(XL, sx);
            if (own && edge_row && keep_sum && g.sum && act) edge_out(famL, y, sx);
        }
        __builtin_amdgcn_sched_barrier(0);
        {
            off += sc.mp;
            lane_step_pk<CH, KL, false, DEFER>(sc, cg, mc, dead);
            if (own && edge_row && keep_sum && g.sum && act) edge_out(1, y, sc);
        }
        __builtin_amdgcn_sched_barrier(0);
        slot_read(XR, c + 1, sx);
        __syncthreads();
        {
            off += sx.mp;
            lane_step_pk<CH, KL, false, DEFER>(sx, cg, mc, dead);
            if (!wave_act) lane_guard<CH>(sx, act);
            slot_write(XR, sx);
            if (own && edge_row && keep_sum && g.sum && act) edge_out(famR, y, sx);
        }
        if (!own || !act) continue;
        if (edge_row) {
            if (DIR == 1 && keep_sum && g.sum) edge_sum(y);   // y == 0
            continue;
        }
        if (!pc) continue;
        // acc = (L_cols + L_FL + L_FR) - 3 M (- the pre-step minima) from the new states of this column,
        // one chunk at a time (scheduling barriers keep the compiler from
        // hoisting every chunk's LDS reads at once)
        auto chunk_sum = [&](auto qc) __attribute__((always_inline)) {
            constexpr int q = decltype(qc)::value;
            const uint2* wl = (const uint2*)XL + (c + 1) * KL + cg.lid + 4 * q * GS;
            const uint2* wr = (const uint2*)XR + (c + 1) * KL + cg.lid + 4 * q * GS;
            uint32_t lv[8], rv[8];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const uint2 a = wl[k * GS], b = wr[k * GS];
                lv[2 * k] = a.x; lv[2 * k + 1] = a.y;
                rv[2 * k] = b.x; rv[2 * k + 1] = b.y;
            }
            uint32_t mq[8];
            unpack16(mc[q], mq);
            const us2 o2 = DEFER ? U2(dup16(off)) : (us2){0, 0};
            // going down: sum - 3 M - off; going up S's 8 M is folded in: sum + 5 M - off
            // (u16 arithmetic modulo 2^16; the true results are below 2^16)
#pragma unroll
            for (int k = 0; k < 8; k++) {
                if (DIR == 0) {
                    const us2 m3 = U2(mq[k]) * (us2){3, 3} + o2;
                    acc[q][k] = W2(U2(sc.L[8 * q + k]) + U2(lv[k]) + U2(rv[k]) - m3);
                } else {
                    const us2 m5 = U2(mq[k]) * (us2){5, 5} - o2;
                    acc[q][k] = W2(U2(sc.L[8 * q + k]) + U2(lv[k]) + U2(rv[k]) + m5);
                }
            }
        };
        if (DIR == 0) {
            static_for<0, CH>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if (!FULL && cg.lid * CH + q >= cg.nq) return;
                chunk_sum(qc);
                uint32_t pr[8];
                const uint4 rw = make_uint4(ob[q].x + ob[CH + q].x, ob[q].y + ob[CH + q].y, ob[q].z + ob[CH + q].z,
                                            ob[q].w + ob[CH + q].w);   // rows fwd + bwd, bytes <= 200
                unpack16(rw, pr);
#pragma unroll
                for (int k = 0; k < 8; k++) acc[q][k] = W2(U2(acc[q][k]) + U2(pr[k]));
                ST16(pd_at(y, q, 0), make_uint4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]));
                ST16(pd_at(y, q, 1), make_uint4(acc[q][4], acc[q][5], acc[q][6], acc[q][7]));
                __builtin_amdgcn_sched_barrier(0);
            });
            continue;
        }
        // upward pass: S and the winner (hpp:283), as k_fam's final sweep
        const int oy = y - D, ox = x - D;
        const bool crop = oy >= 0 && oy < g.H && ox >= 0 && ox < g.W;
        if (!(crop || keep_sum)) continue;
        uint32_t key = 0xFFFFFFFFu;
        const uint32_t m512 = 512u;
        static_for<0, CH>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int qg = cg.lid * CH + q;
            chunk_sum(qc);
            const uint32_t pw[8] = {ob[2 * q].x, ob[2 * q].y, ob[2 * q].z, ob[2 * q].w,
                                    ob[2 * q + 1].x, ob[2 * q + 1].y, ob[2 * q + 1].z, ob[2 * q + 1].w};
#pragma unroll
            for (int k = 0; k < 8; k++) acc[q][k] = W2(U2(pw[k]) + U2(acc[q][k]));   // S pairs
            if (!FULL && qg >= cg.nq) return;
            if (keep_sum && g.sum) {
                uint16_t* sp = g.sum + cell + 16 * q;
                ST16(sp, make_uint4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]));
                ST16(sp + 8, make_uint4(acc[q][4], acc[q][5], acc[q][6], acc[q][7]));
            }
            // keys (S << 9) | (d - the lane's first d): compile-time low bits, so
            // no per-pair lane constants are kept across rows; the lane's base
            // is added after the minimum (d < 512, no carry into S).  In the
            // half-filled last chunk (D % 16 == 8) pairs 4..7 are padding.
            const bool tail = cg.half_tail && qg == cg.nq - 1;
            static_for<0, 8>([&](auto kc) {
                constexpr int k = decltype(kc)::value;
                const uint32_t k0 = vs_key<16 * q + 2 * k, false>(acc[q][k], m512);
                const uint32_t k1 = vs_key<16 * q + 2 * k + 1, true>(acc[q][k], m512);
                const uint32_t kk = min(k0, k1);
                key = (k < 4 || !tail) ? min(key, kk) : key;
            });
            __builtin_amdgcn_sched_barrier(0);
        });
        if (!crop) continue;
        if (key != 0xFFFFFFFFu) key += (uint32_t)(cg.lid * CH * 16);   // lanes of padding only keep ~0
        key = group_min<KL>(key);   // the four lanes of a column share crop
        const uint32_t best = key & 511u;
        float delta = 0.0f;
        if (want_sub) {
            // S at best - 1 and best + 1 from the lanes holding them
            auto held = [&](int d) __attribute__((always_inline)) {
                const int j = (d >> 1) - cg.lid * NP;
                const uint32_t w = vs_pick<CH>(acc, j);
                const uint32_t v = (d & 1) ? (w >> 16) : (w & 0xFFFFu);
                return (d >= 0 && j >= 0 && j < NP) ? v : 0u;
            };
            const int sm = (int)group_max<KL>(held((int)best - 1));
            const int sp = (int)group_max<KL>(held((int)best + 1));
            if (best > 0 && (int)best < D - 1) {
                const int s0 = (int)(key >> 9);
                const int den = sm - 2 * s0 + sp;
                if (den > 0) delta = (float)(sm - sp) / (2.0f * (float)den);
            }
        }
        if (cg.lid == 0) {
            const size_t o = (size_t)oy * g.W + ox;
            if (outm) G(outm)[o] = (uint16_t)min(best * 255u, 65535u);
            if (mode == 0 && g.disp) G(g.disp)[o] = (uint16_t)best;
            if (want_sub) G(g.subpix)[o] = (float)best + delta;
        }
    }
    // states at the end of the segment, for the next launch
    if (t0 + nrows < Hp && own && act) {
        const int pout = seg & 1;
        uint4 mr[CH];
        load_m(t0 + nrows - 1, mr);
        st_store(sc, sg_col(pout, 1), mr);
        slot_read(XL, c, sx);
        st_store(sx, sg_col(pout, famL), mr);
        slot_read(XR, c, sx);
        st_store(sx, sg_col(pout, famR), mr);
    }
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
// SGM launches: (CH chunks per lane, KL lanes per chain) instantiations
#define SISTER_FAM_LAUNCH(CHV, KLV, FAM, ...) hipLaunchKernelGGL((k_fam<CHV, KLV, FAM>), __VA_ARGS__)
#define SISTER_FAM_KL4(CHV, FAM, ...)                                                            \
    switch (CHV) {                                                                               \
    case 1: SISTER_FAM_LAUNCH(1, 4, FAM, __VA_ARGS__); break;                                    \
    case 2: SISTER_FAM_LAUNCH(2, 4, FAM, __VA_ARGS__); break;                                    \
    case 3: SISTER_FAM_LAUNCH(3, 4, FAM, __VA_ARGS__); break;                                    \
    default: SISTER_FAM_LAUNCH(4, 4, FAM, __VA_ARGS__); break;                                   \
    }
#define SISTER_FAM_KL8(CHV, FAM, ...)                                                            \
    switch (CHV) {                                                                               \
    case 1: SISTER_FAM_LAUNCH(1, 8, FAM, __VA_ARGS__); break;                                    \
    case 2: SISTER_FAM_LAUNCH(2, 8, FAM, __VA_ARGS__); break;                                    \
    case 3: SISTER_FAM_LAUNCH(3, 8, FAM, __VA_ARGS__); break;                                    \
    default: SISTER_FAM_LAUNCH(4, 8, FAM, __VA_ARGS__); break;                                   \
    }
#define SISTER_FAM_KL16(CHV, FAM, ...)                                                           \
    switch (CHV) {                                                                               \
    case 1: SISTER_FAM_LAUNCH(1, 16, FAM, __VA_ARGS__); break;                                   \
    default: SISTER_FAM_LAUNCH(2, 16, FAM, __VA_ARGS__); break;                                  \
    }
#define SISTER_FAM(G, FAM, ...)                                                                  \
    do {                                                                                         \
        if ((G).KL[FAM] == 16) { SISTER_FAM_KL16((G).CH[FAM], FAM, __VA_ARGS__); }               \
        else if ((G).KL[FAM] == 8) { SISTER_FAM_KL8((G).CH[FAM], FAM, __VA_ARGS__); }            \
        else { SISTER_FAM_KL4((G).CH[FAM], FAM, __VA_ARGS__); }                                  \
    } while (0)

int sgm_pick_kl(long cols_chains, int fam, int D)
{
    // SISTER_SGM_LANES (4, 8 or 16) overrides the choice, for tuning; 4 lanes
    // hold at most 256 disparities (beyond that the state would spill)
    static const int forced = [] {
        const char* e = getenv("SISTER_SGM_LANES");
        const int v = e ? atoi(e) : 0;
        return (v == 4 || v == 8 || v == 16) ? v : 0;
    }();
    if (forced) return (forced == 4 && D > 256) ? 8 : forced;
    // A single rig has few chains: 16 lanes per chain keep the most waves in
    // flight.  In a batch (measured on 16 rigs of 2304x1464x192 cells) rows
    // and columns are still fastest with 16 lanes per chain, the diagonal
    // families, whose final sweep carries the most registers, with 8.
    if (cols_chains * 4 / 64 < 1024) return 16;
    return fam >= 2 ? 8 : 16;
}

int sgm_pick_ch(int D, int KL)
{
    const int k16 = ((D + 15) / 16);
    int ch = (k16 + KL - 1) / KL;
    if (KL == 16) return ch <= 1 ? 1 : 2;
    return ch <= 3 ? ch : 4;   // KL 8 (D <= 512) and KL 4 (D <= 256)
}

int sgm_family_blocks(const RigDesc& g, int KL, int fam)
{
    const int cpw = 64 / KL;
    const int chains = fam == 0 ? g.Hp : (fam == 1 ? g.Wp : g.Wp + g.Hp - 1);
    return (chains + cpw - 1) / cpw;
}

const char* const kKernelNames[kNumKernels] = {
    "census", "wta", "median_prep", "median_serial", "mask", "msum",
    "sgm_rows", "sgm_cols", "sgm_diag", "sgm_anti", "sgm_vsweep",
};

#define TB(id) do { if (t) t->begin(id, st); } while (0)
#define TE(id) do { if (t) t->end(id, st); } while (0)
#define TB_S(id, s) do { if (t) t->begin(id, s); } while (0)
#define TE_S(id, s) do { if (t) t->end(id, s); } while (0)
#define TD(id, n) do { if (t) t->dispatched[id] += (n); } while (0)

void launch_census(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    TB(kKCensus);
    hipLaunchKernelGGL(k_census, dim3(ls.blocks[kStageCensus]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageCensus]);
    TD(kKCensus, 1);
    TE(kKCensus);
}

void launch_wta(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    TB(kKWta);
    if (ls.wta_rows) {
        // frames 0/2 (rows of W' pixels), then frames 1/3 (rows of H')
        const int waves02 = std::min(kWtaRowMaxWaves, (ls.maxWp + 255) / 256);
        const int waves13 = std::min(kWtaRowMaxWaves, (ls.maxHp + 255) / 256);
        hipLaunchKernelGGL(k_wta_row<false>, dim3(ls.blocks[kStageWta]), dim3(64 * waves02),
                           wta_row_lds(ls.maxWp, ls.maxD), st, ls.rigs, ls.nrig, ls.bs[kStageWta]);
        hipLaunchKernelGGL(k_wta_row<true>, dim3(ls.blocks[kStageWtaV]), dim3(64 * waves13),
                           wta_row_lds(ls.maxHp, ls.maxD), st, ls.rigs, ls.nrig, ls.bs[kStageWtaV]);
        TD(kKWta, 2);
    } else {
        size_t lds = (size_t)2 * (256 + ls.maxD - 1) * sizeof(uint64_t);
        hipLaunchKernelGGL(k_wta, dim3(ls.blocks[kStageWta]), dim3(256), lds, st, ls.rigs, ls.nrig, ls.bs[kStageWta]);
        TD(kKWta, 1);
    }
    // the deep padding bands' rows, copied from each band's first row
    hipLaunchKernelGGL(k_wta_fill, dim3(ls.blocks[kStageWtaFill]), dim3(256), 0, st, ls.rigs, ls.nrig,
                       ls.bs[kStageWtaFill]);
    TD(kKWta, 1);
    TE(kKWta);
}

static size_t g_lds_per_block = 0;   // set by kernels_setup
static int g_num_cu = 0;

void launch_median(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    // With enough (rig, map) blocks to spread over the chip the serial kernel
    // sorts each position's six fixed values itself (no prepass pass over
    // memory); with few, each block's row chain is the critical path and the
    // parallel prepass takes the sorting off it.  Plans with frames wider
    // than the raw form takes run every map from the prepass in one launch.
    const size_t lds = (size_t)2 * (ls.maxCols + 1) * sizeof(uint16_t);
    const bool wide = ls.maxCols > 1024 * kMedRawK;
    const bool raw = !wide && ls.nrig * 8 >= std::max(g_num_cu / 4, 1);
    if (!raw) {
        TB(kKMedianPrep);
        hipLaunchKernelGGL(k_median_prep, dim3(ls.blocks[kStageMedianPrep]), dim3(256), 0, st, ls.rigs, ls.nrig,
                           ls.bs[kStageMedianPrep], ls.scratch);
        TD(kKMedianPrep, 1);
        TE(kKMedianPrep);
    }
    TB(kKMedianSerial);
    if (raw) hipLaunchKernelGGL(k_median_serial<0>, dim3(ls.nrig * 8), dim3(1024), lds, st, ls.rigs, ls.nrig, ls.scratch);
    else if (wide) hipLaunchKernelGGL(k_median_serial<2>, dim3(ls.nrig * 8), dim3(1024), lds, st, ls.rigs, ls.nrig, ls.scratch);
    else hipLaunchKernelGGL(k_median_serial<1>, dim3(ls.nrig * 8), dim3(1024), lds, st, ls.rigs, ls.nrig, ls.scratch);
    TD(kKMedianSerial, 1);
    TE(kKMedianSerial);
}

void launch_mask(const LaunchSet& ls, int fmask, hipStream_t st, KernelTimer* t)
{
    TB(kKMask);
    hipLaunchKernelGGL(k_mask, dim3(ls.blocks[kStageMask]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageMask], fmask);
    TD(kKMask, 1);
    TE(kKMask);
}

size_t msum_lds(int D)
{
    return (size_t)(2 * kWinPad + 2 * (kRun + D - 1) + 2 * kRun) * 8 + kRun;
}

int msum_blocks(const RigDesc& g, int side)
{
    if (side == 2) return ((g.Wp + kMT - 1) / kMT) * ((g.Hp + kMT - 1) / kMT);
    return side == 0 ? g.Hp * ((g.Wp + kRun - 1) / kRun) : g.Wp * ((g.Hp + kRun - 1) / kRun);
}

// at 5 groups (D > 256) the tile's windows need 184,576 bytes of LDS, over
// the 160 KB a block can hold: those disparity counts take the two-launch form
constexpr int kMsumTileMaxJ = 4;

int kernels_setup(int device)
{
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return -1;
    const int lim = v;
    if (hipDeviceGetAttribute(&g_num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return -1;
    const void* fns[kMsumTileMaxJ] = {(const void*)k_msum_tile<1>, (const void*)k_msum_tile<2>,
                                      (const void*)k_msum_tile<3>, (const void*)k_msum_tile<4>};
    for (const void* f : fns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) {
            (void)hipGetLastError();
            g_lds_per_block = 0;   // fused cost sum off: the two-launch form runs
            return 0;
        }
    g_lds_per_block = (size_t)lim;
#define SISTER_VS_FNS(K, F)                                                         \
    (const void*)k_vsweep<1, 0, K, F, 4>, (const void*)k_vsweep<2, 0, K, F, 4>,     \
        (const void*)k_vsweep<3, 0, K, F, 4>, (const void*)k_vsweep<1, 1, K, F, 4>, \
        (const void*)k_vsweep<2, 1, K, F, 4>, (const void*)k_vsweep<3, 1, K, F, 4>, \
        (const void*)k_vsweep<2, 0, K, F, 8>, (const void*)k_vsweep<3, 0, K, F, 8>, \
        (const void*)k_vsweep<2, 1, K, F, 8>, (const void*)k_vsweep<3, 1, K, F, 8>
    for (const void* f : {(const void*)k_wta_row<false>, (const void*)k_wta_row<true>})
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) {
            (void)hipGetLastError();
            g_lds_per_block = 0;
            return 0;
        }
    const void* vfns[30] = {SISTER_VS_FNS(false, false), SISTER_VS_FNS(false, true), SISTER_VS_FNS(true, false)};
#undef SISTER_VS_FNS
    for (const void* f : vfns)
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lim) != hipSuccess) {
            (void)hipGetLastError();
            g_lds_per_block = 0;
            return 0;
        }
    return 0;
}

// SISTER_MSUM_SPLIT=1 keeps the two-launch cost sum everywhere (comparison runs)
static bool msum_split_forced()
{
    static const int v = [] {
        const char* e = getenv("SISTER_MSUM_SPLIT");
        return e && atoi(e) != 0 ? 1 : 0;
    }();
    return v != 0;
}

static bool msum_tile_fits(const DGroup& G)
{
    return (G.Dp + 63) / 64 <= kMsumTileMaxJ && msum_tile_lds(G.Dp) <= g_lds_per_block;
}

static void launch_msum_tile(const DGroup& G, int fuse, hipStream_t st)
{
    const size_t lds = msum_tile_lds(G.Dp);
    const int ntiles = G.blocks[kGMsumTile];
    const dim3 grid(std::min(ntiles, std::max(g_num_cu, 1))), block(1024);   // persistent, one per CU
    const int* bs = G.bs[kGMsumTile];
    switch ((G.Dp + 63) / 64) {
    case 1: hipLaunchKernelGGL(k_msum_tile<1>, grid, block, lds, st, G.rigs, G.nrig, bs, ntiles, fuse); break;
    case 2: hipLaunchKernelGGL(k_msum_tile<2>, grid, block, lds, st, G.rigs, G.nrig, bs, ntiles, fuse); break;
    case 3: hipLaunchKernelGGL(k_msum_tile<3>, grid, block, lds, st, G.rigs, G.nrig, bs, ntiles, fuse); break;
    default: hipLaunchKernelGGL(k_msum_tile<4>, grid, block, lds, st, G.rigs, G.nrig, bs, ntiles, fuse); break;
    }
}

template <int SIDE>
static void launch_msum_side(const DGroup& G, int fuse, hipStream_t st)
{
    const size_t lds = msum_lds(G.D);
    const int k = kGMsumRows + SIDE;
    const dim3 grid(G.blocks[k]), block(256);
    switch ((G.Dp + 63) / 64) {
    case 1: hipLaunchKernelGGL((k_msum_run<SIDE, 1>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 2: hipLaunchKernelGGL((k_msum_run<SIDE, 2>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 3: hipLaunchKernelGGL((k_msum_run<SIDE, 3>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 4: hipLaunchKernelGGL((k_msum_run<SIDE, 4>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 5: hipLaunchKernelGGL((k_msum_run<SIDE, 5>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 6: hipLaunchKernelGGL((k_msum_run<SIDE, 6>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    case 7: hipLaunchKernelGGL((k_msum_run<SIDE, 7>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    default: hipLaunchKernelGGL((k_msum_run<SIDE, 8>), grid, block, lds, st, G.rigs, G.nrig, G.bs[k], fuse); break;
    }
}

void launch_msum(const LaunchSet& ls, int fuse, hipStream_t st, KernelTimer* t)
{
    TB(kKMsum);
    for (int gi = 0; gi < ls.ngroups; gi++) {
        const DGroup& G = ls.groups[gi];
        if ((fuse & 5) && (fuse & 10) && msum_tile_fits(G) && !G.msum_split && !msum_split_forced()) {
            launch_msum_tile(G, fuse, st);                 // all frames, M written once
            TD(kKMsum, 1);
            continue;
        }
        if (fuse & 5) launch_msum_side<0>(G, fuse, st);    // frames 0/180 write M (zero when alone)
        if (fuse & 10) launch_msum_side<1>(G, fuse, st);   // frames 90/270 add (or write when alone)
        TD(kKMsum, ((fuse & 5) ? 1 : 0) + ((fuse & 10) ? 1 : 0));
    }
    TE(kKMsum);
}

// lanes per column of the fused sweep for D (0: beyond its reach, D > 384)
static int vs_kl(int D)
{
    const int k16 = (D + 15) / 16;
    return (k16 + 3) / 4 <= 3 ? 4 : (k16 + 7) / 8 <= 3 ? 8 : 0;
}
static int vs_ch(int D) { const int kl = vs_kl(D); return kl ? ((D + 15) / 16 + kl - 1) / kl : 0; }

// the fused sweep can keep P over the rows family's bytes (no volumes of its own)
bool vsweep_in_place(int D)
{
    const int kl = vs_kl(D);
    return kl && D == 16 * kl * vs_ch(D);
}

static size_t vs_lds(int D, int vcols)
{
    const int kl = vs_kl(D);
    const size_t nt = vcols + 2 * kVGhost;
    return 2 * ((size_t)8 * vs_ch(D) * vs_gs((int)nt, kl) + nt + 2) * sizeof(uint32_t);
}

int vsweep_pick_cols(long cols, int D, int choice)
{
    // choice: 0 separate families, 1 fused sweep where D allows, -1 automatic;
    // SISTER_VSWEEP=0/1 in the environment overrides the automatic choice
    static const int env = [] {
        const char* e = getenv("SISTER_VSWEEP");
        return e ? atoi(e) : -1;
    }();
    const int forced = choice >= 0 ? choice : env;
    // up to 3 chunks of 16 per lane (beyond, the state spills): 4 lanes per
    // column up to D = 192, 8 up to D = 384
    if (forced == 0 || vs_kl(D) == 0) return 0;
    const int vcols = vs_cols(vs_kl(D));
    if (vs_lds(D, vcols) > g_lds_per_block) return 0;
    // enough blocks to fill the chip (few rigs keep the per-chain families,
    // which spread a single rig over every CU)
    if (forced != 1 && (cols + vcols - 1) / vcols < std::max(g_num_cu / 2, 1)) return 0;
    return vcols;
}

int vsweep_blocks(const RigDesc& g, int vcols)
{
    return vcols > 0 ? (g.Wp + vcols - 1) / vcols : 0;
}

static int launch_vsweep(const DGroup& G, int mode, int keep_sum, hipStream_t st)
{
    const int kl = vs_kl(G.D);
    const int nt = G.vcols + 2 * kVGhost;
    const size_t lds = vs_lds(G.D, G.vcols);
    const dim3 grid(G.blocks[kGVSweep]), block(kl * nt);
    const int* bs = G.bs[kGVSweep];
    const int nseg = (G.maxHp + kVGhost - 1) / kVGhost;
    // 1..3 chunks with 4 lanes per column, 2..3 with 8 (vsweep_pick_cols
    // gives no columns beyond)
    const int ch = vs_ch(G.D);
    const bool full = G.D == G.Dp && G.Dp == 16 * kl * ch;
    for (int dir = 0; dir < 2; dir++)
        for (int seg = 0; seg < nseg; seg++) {
#define SISTER_VS(CHV, DIRV, KLV)                                                                              \
    do {                                                                                                       \
        if (keep_sum)                                                                                          \
            hipLaunchKernelGGL((k_vsweep<CHV, DIRV, true, false, KLV>), grid, block, lds, st, G.rigs, G.nrig,  \
                               bs, G.vcols, seg, mode);                                                        \
        else if (full)                                                                                         \
            hipLaunchKernelGGL((k_vsweep<CHV, DIRV, false, true, KLV>), grid, block, lds, st, G.rigs, G.nrig,  \
                               bs, G.vcols, seg, mode);                                                        \
        else                                                                                                   \
            hipLaunchKernelGGL((k_vsweep<CHV, DIRV, false, false, KLV>), grid, block, lds, st, G.rigs, G.nrig, \
                               bs, G.vcols, seg, mode);                                                        \
    } while (0)
            if (kl == 4) {
                if (dir == 0) {
                    switch (ch) {
                    case 1: SISTER_VS(1, 0, 4); break;
                    case 2: SISTER_VS(2, 0, 4); break;
                    default: SISTER_VS(3, 0, 4); break;
                    }
                } else {
                    switch (ch) {
                    case 1: SISTER_VS(1, 1, 4); break;
                    case 2: SISTER_VS(2, 1, 4); break;
                    default: SISTER_VS(3, 1, 4); break;
                    }
                }
            } else {
                if (dir == 0) {
                    if (ch == 2) SISTER_VS(2, 0, 8);
                    else SISTER_VS(3, 0, 8);
                } else {
                    if (ch == 2) SISTER_VS(2, 1, 8);
                    else SISTER_VS(3, 1, 8);
                }
            }
#undef SISTER_VS
        }
    return 2 * nseg;
}

void launch_sgm(const LaunchSet& ls, int mode, int keep_sum, hipStream_t st, const FamStreams* fs, KernelTimer* t)
{
    bool all_fused = true;
    for (int gi = 0; gi < ls.ngroups; gi++) all_fused = all_fused && ls.groups[gi].vcols > 0;
    if (all_fused) {
        // rows family (its two directions as separate waves: 29.7-30.1 ms per
        // 64-rig step against 31.3-35.3 with both in one wave), then the fused
        // vertical sweep (which reads part[rows])
        for (int gi = 0; gi < ls.ngroups; gi++) {
            const DGroup& G = ls.groups[gi];
            TB_S(kKRows, st);
            SISTER_FAM(G, 0, dim3(2 * G.blocks[0]), dim3(64), 0, st, G.rigs, G.nrig, G.bs[0], mode, keep_sum, 1);
            TD(kKRows, 1);
            TE_S(kKRows, st);
            TB_S(kKVSweep, st);
            const int nl = launch_vsweep(G, mode, keep_sum, st);
            TD(kKVSweep, nl);
            TE_S(kKVSweep, st);
        }
        return;
    }
    // rows on st, columns and diagonals (then the anti-diagonal family's
    // forward sweep) on the side streams, joined before the anti-diagonal
    // family's final sweep
    hipStream_t sc = st, sd = st;
    if (fs) {
        (void)hipEventRecord(fs->fork, st);
        (void)hipStreamWaitEvent(fs->side[0], fs->fork, 0);
        (void)hipStreamWaitEvent(fs->side[1], fs->fork, 0);
        sc = fs->side[0];
        sd = fs->side[1];
    }
    for (int gi = 0; gi < ls.ngroups; gi++) {
        const DGroup& G = ls.groups[gi];
        if (G.vcols > 0) continue;
        TB_S(kKRows, st);
        SISTER_FAM(G, 0, dim3(2 * G.blocks[0]), dim3(64), 0, st, G.rigs, G.nrig, G.bs[0], mode, keep_sum, 1);
        TD(kKRows, 1);
        TE_S(kKRows, st);
        TB_S(kKCols, sc);
        SISTER_FAM(G, 1, dim3(G.blocks[1]), dim3(64), 0, sc, G.rigs, G.nrig, G.bs[1], mode, keep_sum, 0);
        TD(kKCols, 1);
        TE_S(kKCols, sc);
        TB_S(kKDiag, sd);
        SISTER_FAM(G, 2, dim3(G.blocks[2]), dim3(64), 0, sd, G.rigs, G.nrig, G.bs[2], mode, keep_sum, 0);
        TD(kKDiag, 1);
        TE_S(kKDiag, sd);
        // the anti-diagonal family's forward sweep needs M alone: it follows
        // the diagonals on their stream, beside the rows
        TB_S(kKAnti, sd);
        SISTER_FAM(G, 3, dim3(G.blocks[3]), dim3(64), 0, sd, G.rigs, G.nrig, G.bs[3], mode, keep_sum, 1);
        TD(kKAnti, 1);
        TE_S(kKAnti, sd);
    }
    if (fs) {
        (void)hipEventRecord(fs->join[0], sc);
        (void)hipEventRecord(fs->join[1], sd);
        (void)hipStreamWaitEvent(st, fs->join[0], 0);
        (void)hipStreamWaitEvent(st, fs->join[1], 0);
    }
    for (int gi = 0; gi < ls.ngroups; gi++) {
        const DGroup& G = ls.groups[gi];
        if (G.vcols > 0) {
            TB_S(kKRows, st);
            SISTER_FAM(G, 0, dim3(2 * G.blocks[0]), dim3(64), 0, st, G.rigs, G.nrig, G.bs[0], mode, keep_sum, 1);
            TD(kKRows, 1);
            TE_S(kKRows, st);
            TB_S(kKVSweep, st);
            const int nl = launch_vsweep(G, mode, keep_sum, st);
            TD(kKVSweep, nl);
            TE_S(kKVSweep, st);
            continue;
        }
        const size_t lds = (size_t)64 * G.CH[3] * 32;
        TB_S(kKAnti, st);
        SISTER_FAM(G, 3, dim3(G.blocks[3]), dim3(64), lds, st, G.rigs, G.nrig, G.bs[3], mode, keep_sum, 2);
        TD(kKAnti, 1);
        TE_S(kKAnti, st);
    }
}

// per-rig block counts of each stage (the host builds the block-start tables)
bool wta_rows_ok(int maxCols, int maxD)
{
    // SISTER_WTA_ROWS=0 keeps the per-segment kernel (comparison runs)
    static const int env = [] {
        const char* e = getenv("SISTER_WTA_ROWS");
        return e ? atoi(e) : -1;
    }();
    if (env == 0) return false;
    return (maxCols + 255) / 256 <= 4 * kWtaRowMaxWaves && wta_row_lds(maxCols, maxD) <= g_lds_per_block;
}

void stage_blocks(const RigDesc& g, int maxD, bool wta_rows, int out[kNumStages])
{
    (void)maxD;
    const long npx = (long)g.Wp * g.Hp;
    out[kStageCensus] = 8 * ((g.Wp + kCT - 1) / kCT) * ((g.Hp + kCT - 1) / kCT);
    if (wta_rows) {
        out[kStageWta] = 2 * g.Hp;     // rows of frames 0 and 2
        out[kStageWtaV] = 2 * g.Wp;    // rows of frames 1 and 3
    } else {
        out[kStageWta] = 2 * g.Hp * ((g.Wp + 255) / 256) + 2 * g.Wp * ((g.Hp + 255) / 256);
        out[kStageWtaV] = 0;
    }
    out[kStageWtaFill] = 8 * 2 * ((std::max(g.Wp, g.Hp) + kWtaFillCols - 1) / kWtaFillCols);
    out[kStageMedianPrep] = (int)(8 * ((npx + 255) / 256));
    out[kStageMask] = 4 * ((g.Wp + kMT32 - 1) / kMT32) * ((g.Hp + kMT32 - 1) / kMT32);
}

}  // namespace sister
