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
//   accumulateCostsSSE     sgm.cpp:26-455               k_sgm_rows / k_sgm_cols /
//                                                       k_sgm_diag / k_sgm_anti
//   WTALeft on S, convert  hpp:283,111-118              fused into k_sgm_anti
//
// Every kernel takes the rig table plus a block-start table so one launch can
// cover many rigs of different sizes: block b belongs to rig i when
// bs[i] <= b < bs[i+1].
#include <hip/hip_runtime.h>
#include "rig.h"
#include "kernels.h"

namespace sister {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ int find_rig(const int* __restrict__ bs, int n, int b)
{
    int lo = 0, hi = n;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (bs[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
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
// K1 census (census.cpp:38-51), one 16x16 tile of frame pixels per block.
// Test bit for offset (y,x), y=-3..3 outer, x=-4..4 inner, is
// I(r+y, c+x) > I(r-y, c-x), shifted in MSB-first; bit 63 carries bit 0 of
// the previous domain pixel in frame raster order (the accumulator is never
// reset, census.cpp:42), i.e. the (3,4) test of that pixel.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_census(const RigDesc* __restrict__ rigs, int nrig,
                                                const int* __restrict__ bs)
{
    __shared__ uint8_t tile[22][26];
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lb = blockIdx.x - bs[ri];
    const int T = ((g.Wp + 15) / 16) * ((g.Hp + 15) / 16);
    const int unit = lb / T, t = lb % T;
    const int F = unit >> 1;
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const int tilesC = (C + 15) / 16;
    const int r0 = (t / tilesC) * 16, c0 = (t % tilesC) * 16;
    const uint8_t* v = g.view[(unit & 1) ? 1 + F : 0];

    for (int i = threadIdx.x; i < 22 * 25; i += 256) {
        int rr = i / 25, cc = i % 25;
        int r = r0 - 3 + rr, c = c0 - 5 + cc;
        tile[rr][cc] = (r >= 0 && r < R && c >= 0 && c < C) ? (uint8_t)frame_px(g, v, F, r, c) : 0;
    }
    __syncthreads();
    const int r = r0 + (threadIdx.x >> 4), c = c0 + (threadIdx.x & 15);
    if (r >= R || c >= C) return;
    uint64_t code = 0;
    if (r >= 3 && r < R - 3 && c >= 4 && c < C - 4) {
        const int lr = r - r0 + 3, lc = c - c0 + 5;
#pragma unroll
        for (int y = -3; y <= 3; y++)
#pragma unroll
            for (int x = -4; x <= 4; x++)
                code = (code << 1) | (uint64_t)(tile[lr + y][lc + x] > tile[lr - y][lc - x]);
        unsigned carry;
        if (c > 4)
            carry = tile[lr + 3][lc + 3] > tile[lr - 3][lc - 5];          // previous pixel (r, c-1)
        else if (r > 3)
            carry = frame_px(g, v, F, r + 2, C - 1) > frame_px(g, v, F, r - 4, C - 9);  // (r-1, C-5)
        else
            carry = 0;
        code |= (uint64_t)carry << 63;
    }
    g.code[unit][(size_t)r * C + c] = code;
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
    (void)R;
    const int tilesC = (C + 255) / 256;
    const int r = lb / tilesC, c0 = (lb % tilesC) * 256;
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

__device__ __forceinline__ unsigned med9_pre(uint64_t s, unsigned a, unsigned b, unsigned c)
{
    cswap(a, b); cswap(b, c); cswap(a, b);
    const unsigned s1 = (unsigned)(s & 0xFFFF), s2 = (unsigned)((s >> 16) & 0xFFFF);
    const unsigned s3 = (unsigned)((s >> 32) & 0xFFFF), s4 = (unsigned)(s >> 48);
    return min(min(s4, max(s3, a)), min(max(s2, b), max(s1, c)));
}

// One block per (rig, map), 1024 threads; rows in sequence, columns in parallel.
__global__ __launch_bounds__(1024) void k_median_serial(const RigDesc* __restrict__ rigs,
                                                        uint64_t* const* __restrict__ scratch)
{
    extern __shared__ uint16_t lds16[];
    const int ri = blockIdx.x >> 3, m = blockIdx.x & 7;
    const RigDesc& g = rigs[ri];
    const int F = m >> 1;
    const int R = frame_rows(g, F), C = frame_cols(g, F);
    const long npx = (long)R * C;
    const uint16_t* in = g.wta[m];
    uint16_t* out = g.med[m];
    const uint64_t* pre = scratch[ri] + (size_t)m * npx;
    const long qhi = npx - C - 5;
    uint16_t* prev = lds16;
    uint16_t* cur = lds16 + C;
    const int tid = threadIdx.x;
    for (int c = tid; c < C; c += 1024) {
        uint16_t v = in[c];
        prev[c] = v;
        out[c] = v;
    }
    __syncthreads();
    unsigned pp = 0;  // out(r-2, C-1)
    for (int r = 1; r < R; r++) {
        const long base = (long)r * C;
        unsigned ppn = prev[C - 1];
        for (int c = tid; c < C - 1; c += 1024) {
            const long q = base + c;
            unsigned v;
            if (q == C) v = 0;
            else if (q > qhi) v = in[q];
            else {
                unsigned a = (c == 0) ? pp : prev[c - 1];
                v = med9_pre(pre[q], a, prev[c], prev[c + 1]);
            }
            cur[c] = (uint16_t)v;
        }
        __syncthreads();
        if (tid == 0) {
            const long q = base + C - 1;
            unsigned v;
            if (q > qhi) v = in[q];
            else {
                v = med9_pre(pre[q], prev[C - 2], prev[C - 1], cur[0]);
            }
            cur[C - 1] = (uint16_t)v;
        }
        __syncthreads();
        for (int c = tid; c < C; c += 1024) out[base + c] = cur[c];
        pp = ppn;
        uint16_t* t = prev; prev = cur; cur = t;
    }
}

// ---------------------------------------------------------------------------
// K4 LR check + mask (postprocess.cpp:318-341, hpp:201-206 etc.): frame pixel
// (r,c) keeps L when 0 <= L <= c and |L - R[c-L]| <= 5; mask = kept && L > 0
// && c >= D, written at the padded original coordinates (hpp:216,235-236,250-251).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mask(const RigDesc* __restrict__ rigs, int nrig,
                                              const int* __restrict__ bs, int fmask)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const long npx = (long)g.Wp * g.Hp;
    const long perF = (npx + 255) / 256;
    const long lb = blockIdx.x - bs[ri];
    const int F = (int)(lb / perF);
    const long q = (lb % perF) * 256 + threadIdx.x;
    if (q >= npx || !(fmask & (1 << F))) return;
    const int C = frame_cols(g, F);
    const int r = (int)(q / C), c = (int)(q % C);
    const int L = g.med[2 * F][q];
    unsigned ok = 0;
    if (c >= g.D && L >= 1 && L <= c) {
        int Rv = g.med[2 * F + 1][q - L];
        ok = abs(L - Rv) <= kLrThreshold;
    }
    int y, x;
    frame_to_orig(g, F, r, c, y, x);
    g.mask[F][(size_t)y * g.Wp + x] = (uint8_t)ok;
}

// ---------------------------------------------------------------------------
// K5 mask-weighted cost sum (hpp:255-277).  M(y,x,d) = sum over used frames F
// with mask_F(y,x) = 1 of pop(CC_F[r][c] ^ CN_F[r][c-d]), (r,c) the frame
// pixel of (y,x).  A set mask implies c >= D > d, so every term is a plain
// popcount (<= 63) and M <= 252 fits a byte.  One wave per pixel, lanes over d.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_msum(const RigDesc* __restrict__ rigs, int nrig,
                                              const int* __restrict__ bs, int fuse)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x & 63;
    const long p = (long)(blockIdx.x - bs[ri]) * 4 + (threadIdx.x >> 6);
    const long npx = (long)g.Wp * g.Hp;
    if (p >= npx) return;
    const int D = g.D;
    const int y = (int)(p / g.Wp), x = (int)(p % g.Wp);
    unsigned acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int F = 0; F < 4; F++) {
        if (!(fuse & (1 << F)) || !g.mask[F][p]) continue;
        int r, c;
        orig_to_frame(g, F, y, x, r, c);
        if (c < D) continue;   // a set mask implies c >= D; never index before the row
        const int C = frame_cols(g, F);
        const uint64_t cc = g.code[2 * F][(size_t)r * C + c];
        const uint64_t* cn = g.code[2 * F + 1] + (size_t)r * C + c;
#pragma unroll
        for (int j = 0; j < 8; j++) {
            int d = lane + 64 * j;
            if (d < D) acc[j] += __popcll(cc ^ cn[-d]);
        }
    }
    uint8_t* o = g.msum + p * D;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        int d = lane + 64 * j;
        if (d < D) o[d] = (uint8_t)acc[j];
    }
}

// ---------------------------------------------------------------------------
// K6 SGM (sgm.cpp:26-455).  S is the sum of eight path costs; each path is an
// independent chain, so the four chain families (rows, columns, diagonals,
// anti-diagonals) are swept by separate kernels, one wave per chain and both
// directions per wave, lanes over d (d = lane + 64 j).
//
// Standard step (rows >= 1 of a pass, all paths; sgm.cpp:278-383):
//   L(d) = M(d) + min(L'(d), min(L'(d-1), L'(d+1)) + P1, minL' + P2) - minL'
// with L'(-1) = L'(D) = guard.  Path values stay <= M + P2 <= 352 and S <=
// 8*352, so the reference's saturating u16 arithmetic never saturates and
// plain u32 arithmetic gives the same numbers.
// Pass structure reproduced exactly (sgm.cpp:89-207):
//   - first line of each pass (row 0 down, row H'-1 up): paths r1,r2,r3 only
//     take the state c~ = (M == 255 ? 0 : M) and add nothing to S; the
//     horizontal path r0 runs on c~ with its result truncated to 8 bits
//     (uint8 saturate_cast) and is the only contribution on that row;
//   - r0 restarts from L' = 0, minL' = 0 on every other row (L = M at the
//     first pixel);
//   - r1 / r3 entering from outside the left/right border see L' = guard and
//     minL' = 0, giving L = M + P2.
// ---------------------------------------------------------------------------
template <int J>
struct Vec { unsigned v[J]; };

template <int J>
__device__ __forceinline__ void sgm_step(const Vec<J>& Lp, unsigned mp, const Vec<J>& c, Vec<J>& Ln,
                                         unsigned& mn, int D, int lane, bool trunc8)
{
    unsigned lmin = 0xFFFFFFFFu;
    const unsigned p2 = mp + kP2;
#pragma unroll
    for (int j = 0; j < J; j++) {
        unsigned fl = (j == 0) ? kGuard : (unsigned)__builtin_amdgcn_readlane((int)Lp.v[j > 0 ? j - 1 : 0], 63);
        unsigned fr = (j == J - 1) ? kGuard : (unsigned)__builtin_amdgcn_readlane((int)Lp.v[j + 1 < J ? j + 1 : j], 0);
        unsigned nb = min(shr1(Lp.v[j], fl), shl1(Lp.v[j], fr)) + kP1;
        unsigned t = min(min(nb, Lp.v[j]), p2);
        unsigned v = c.v[j] + t - mp;
        if (trunc8) v &= 0xFFu;
        v = (lane + 64 * j < D) ? v : kGuard;
        Ln.v[j] = v;
        lmin = min(lmin, v);
    }
    mn = wave_min(lmin);
}

template <int J>
__device__ __forceinline__ void load_cost(Vec<J>& c, const uint8_t* cell, int D, int lane, bool first_line)
{
#pragma unroll
    for (int j = 0; j < J; j++) {
        int d = lane + 64 * j;
        unsigned v = (d < D) ? cell[d] : 0u;
        if (first_line && v == 255u) v = 0u;   // paramInvalidDispCost (sgm.cpp:109-110,146-147)
        c.v[j] = v;
    }
}

template <int J>
__device__ __forceinline__ void init_state(Vec<J>& L, unsigned& m, const Vec<J>& c, int D, int lane, unsigned add)
{
    unsigned lmin = 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < J; j++) {
        unsigned v = (lane + 64 * j < D) ? c.v[j] + add : kGuard;
        L.v[j] = v;
        lmin = min(lmin, v);
    }
    m = wave_min(lmin);
}

template <int J>
__device__ __forceinline__ void load_sum(Vec<J>& s, const uint16_t* cell, int D, int lane)
{
#pragma unroll
    for (int j = 0; j < J; j++) {
        int d = lane + 64 * j;
        s.v[j] = (d < D) ? cell[d] : 0u;
    }
}

template <int J>
__device__ __forceinline__ void store_sum(uint16_t* cell, const Vec<J>& s, int D, int lane)
{
#pragma unroll
    for (int j = 0; j < J; j++) {
        int d = lane + 64 * j;
        if (d < D) cell[d] = (uint16_t)s.v[j];
    }
}

template <int J>
__device__ __forceinline__ void acc_sum(uint16_t* cell, const Vec<J>& L, int D, int lane)
{
    Vec<J> s;
    load_sum<J>(s, cell, D, lane);
#pragma unroll
    for (int j = 0; j < J; j++) s.v[j] += L.v[j];
    store_sum<J>(cell, s, D, lane);
}

// one chain walked cell by cell: visit(i) -> (y, x) of the i-th cell
// kind: 0 = first line (state c~, no contribution), 1 = guard entry (M + P2,
// contributes), 2 = restart of r0 (L = M, contributes), 3 = first line of r0
// (c~, contributes, later steps truncated)
template <int J, class Pos, class Emit>
__device__ __forceinline__ void run_chain(const RigDesc& g, int n, int kind, Pos pos, Emit emit, int lane)
{
    const int D = g.D;
    const size_t rowc = (size_t)g.Wp * D;
    Vec<J> L, c;
    unsigned m;
    int y, x;
    pos(0, y, x);
    load_cost<J>(c, g.msum + y * rowc + (size_t)x * D, D, lane, kind == 0 || kind == 3);
    init_state<J>(L, m, c, D, lane, kind == 1 ? (unsigned)kP2 : 0u);
    if (kind != 0) emit(y, x, L);
    const bool trunc = (kind == 3);   // the whole chain lies on the first line
    for (int i = 1; i < n; i++) {
        pos(i, y, x);
        load_cost<J>(c, g.msum + y * rowc + (size_t)x * D, D, lane, trunc);
        Vec<J> Ln;
        unsigned mn;
        sgm_step<J>(L, m, c, Ln, mn, D, lane, trunc);
        L = Ln;
        m = mn;
        emit(y, x, L);
    }
}

// r0 (horizontal): one wave per row; forward sweep writes S, backward adds.
template <int J>
__global__ __launch_bounds__(256) void k_sgm_rows(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x & 63;
    const int y = (blockIdx.x - bs[ri]) * 4 + (threadIdx.x >> 6);
    if (y >= g.Hp) return;
    const int D = g.D, Wp = g.Wp;
    uint16_t* srow = g.sum + (size_t)y * Wp * D;
    // pass 0 direction: left to right; row 0 is the pass-0 first line
    run_chain<J>(g, Wp, y == 0 ? 3 : 2,
        [&](int i, int& yy, int& xx) { yy = y; xx = i; },
        [&](int yy, int xx, const Vec<J>& L) { store_sum<J>(srow + (size_t)xx * D, L, D, lane); }, lane);
    // pass 1 direction: right to left; row H'-1 is the pass-1 first line
    run_chain<J>(g, Wp, y == g.Hp - 1 ? 3 : 2,
        [&](int i, int& yy, int& xx) { yy = y; xx = Wp - 1 - i; },
        [&](int yy, int xx, const Vec<J>& L) { acc_sum<J>(srow + (size_t)xx * D, L, D, lane); }, lane);
}

// r2 (vertical): one wave per column
template <int J>
__global__ __launch_bounds__(256) void k_sgm_cols(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x & 63;
    const int x = (blockIdx.x - bs[ri]) * 4 + (threadIdx.x >> 6);
    if (x >= g.Wp) return;
    const int D = g.D, Hp = g.Hp;
    const size_t rowc = (size_t)g.Wp * D;
    auto emit = [&](int yy, int xx, const Vec<J>& L) { acc_sum<J>(g.sum + yy * rowc + (size_t)xx * D, L, D, lane); };
    run_chain<J>(g, Hp, 0, [&](int i, int& yy, int& xx) { yy = i; xx = x; }, emit, lane);
    run_chain<J>(g, Hp, 0, [&](int i, int& yy, int& xx) { yy = Hp - 1 - i; xx = x; }, emit, lane);
}

// r1 (down-right, pass 0) and its pass-1 mirror (up-left): chain k = x - y
template <int J>
__global__ __launch_bounds__(256) void k_sgm_diag(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x & 63;
    const int kk = (blockIdx.x - bs[ri]) * 4 + (threadIdx.x >> 6);
    const int Wp = g.Wp, Hp = g.Hp, D = g.D;
    if (kk >= Wp + Hp - 1) return;
    const int k = kk - (Hp - 1);                 // x - y
    const int y0 = max(0, -k);
    const int y1 = min(Hp - 1, Wp - 1 - k);
    const int n = y1 - y0 + 1;
    const size_t rowc = (size_t)Wp * D;
    auto emit = [&](int yy, int xx, const Vec<J>& L) { acc_sum<J>(g.sum + yy * rowc + (size_t)xx * D, L, D, lane); };
    // down: enters at row 0 (first line) or at column 0 (left border guard)
    run_chain<J>(g, n, y0 == 0 ? 0 : 1, [&](int i, int& yy, int& xx) { yy = y0 + i; xx = yy + k; }, emit, lane);
    // up: enters at row H'-1 (first line) or at column W'-1 (right border guard)
    run_chain<J>(g, n, y1 == Hp - 1 ? 0 : 1, [&](int i, int& yy, int& xx) { yy = y1 - i; xx = yy + k; }, emit, lane);
}

// r3 (down-left, pass 0) and its pass-1 mirror (up-right): chain s = x + y.
// The upward sweep is the last contribution to every cell it visits, so it
// also takes the final winner (hpp:283) for cells inside the crop and writes
// the API outputs (hpp:111-118) plus the sub-pixel refinement.
template <int J>
__global__ __launch_bounds__(256) void k_sgm_anti(const RigDesc* __restrict__ rigs, int nrig,
                                                  const int* __restrict__ bs, int mode, int keep_sum)
{
    const int ri = find_rig(bs, nrig, blockIdx.x);
    const RigDesc& g = rigs[ri];
    const int lane = threadIdx.x & 63;
    const int s = (blockIdx.x - bs[ri]) * 4 + (threadIdx.x >> 6);
    const int Wp = g.Wp, Hp = g.Hp, D = g.D;
    if (s >= Wp + Hp - 1) return;
    const int y0 = max(0, s - (Wp - 1));
    const int y1 = min(Hp - 1, s);
    const int n = y1 - y0 + 1;
    const size_t rowc = (size_t)Wp * D;
    auto emit = [&](int yy, int xx, const Vec<J>& L) { acc_sum<J>(g.sum + yy * rowc + (size_t)xx * D, L, D, lane); };
    // down: from (y-1, x+1); enters at row 0 or at column W'-1
    run_chain<J>(g, n, y0 == 0 ? 0 : 1, [&](int i, int& yy, int& xx) { yy = y0 + i; xx = s - yy; }, emit, lane);
    // up: from (y+1, x-1); enters at row H'-1 or at column 0
    uint16_t* outm = g.out[mode];
    auto final_emit = [&](int yy, int xx, const Vec<J>& L) {
        uint16_t* cell = g.sum + yy * rowc + (size_t)xx * D;
        Vec<J> S;
        load_sum<J>(S, cell, D, lane);
#pragma unroll
        for (int j = 0; j < J; j++) S.v[j] += L.v[j];
        const int oy = yy - D, ox = xx - D;
        const bool crop = !(oy < 0 || oy >= g.H || ox < 0 || ox >= g.W);
        if (!crop || keep_sum) store_sum<J>(cell, S, D, lane);
        if (!crop) return;
        // WTALeft over the full range (x >= D > D-1), lowest index on ties
        unsigned key = 0xFFFFFFFFu;
#pragma unroll
        for (int j = 0; j < J; j++) {
            int d = lane + 64 * j;
            if (d < D) key = min(key, (S.v[j] << 9) | (unsigned)d);
        }
        key = wave_min(key);
        const unsigned best = key & 511u;
        const size_t o = (size_t)oy * g.W + ox;
        if (lane == 0) {
            if (outm) outm[o] = (uint16_t)min(best * 255u, 65535u);
            if (mode == 0 && g.disp) g.disp[o] = (uint16_t)best;
        }
        if (mode == 0 && g.subpix) {
            // parabola through S(d*-1), S(d*), S(d*+1) (build-defined refinement)
            float delta = 0.0f;
            if (best > 0 && (int)best < D - 1) {
                unsigned sv[3];
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const unsigned dd = best - 1 + t;
                    const unsigned jj = dd >> 6, ll = dd & 63;
                    unsigned val = 0;
#pragma unroll
                    for (int j = 0; j < J; j++)
                        if ((unsigned)j == jj) val = (unsigned)__builtin_amdgcn_readlane((int)S.v[j], (int)ll);
                    sv[t] = val;
                }
                const int den = (int)sv[0] - 2 * (int)sv[1] + (int)sv[2];
                if (den > 0) delta = (float)((int)sv[0] - (int)sv[2]) / (2.0f * (float)den);
            }
            if (lane == 0) g.subpix[o] = (float)best + delta;
        }
    };
    run_chain<J>(g, n, y1 == Hp - 1 ? 0 : 1, [&](int i, int& yy, int& xx) { yy = y1 - i; xx = s - yy; }, final_emit, lane);
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
#define SISTER_J_DISPATCH(J, KERNEL, ...)                                                        \
    switch (J) {                                                                                 \
    case 1: hipLaunchKernelGGL(KERNEL<1>, __VA_ARGS__); break;                                   \
    case 2: hipLaunchKernelGGL(KERNEL<2>, __VA_ARGS__); break;                                   \
    case 3: hipLaunchKernelGGL(KERNEL<3>, __VA_ARGS__); break;                                   \
    case 4: hipLaunchKernelGGL(KERNEL<4>, __VA_ARGS__); break;                                   \
    case 5: hipLaunchKernelGGL(KERNEL<5>, __VA_ARGS__); break;                                   \
    case 6: hipLaunchKernelGGL(KERNEL<6>, __VA_ARGS__); break;                                   \
    case 7: hipLaunchKernelGGL(KERNEL<7>, __VA_ARGS__); break;                                   \
    default: hipLaunchKernelGGL(KERNEL<8>, __VA_ARGS__); break;                                  \
    }

const char* const kKernelNames[kNumKernels] = {
    "census", "wta", "median_prep", "median_serial", "mask", "msum",
    "sgm_rows", "sgm_cols", "sgm_diag", "sgm_anti",
};

#define TB(id) do { if (t) t->begin(id, st); } while (0)
#define TE(id) do { if (t) t->end(id, st); } while (0)

void launch_census(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    TB(kKCensus);
    hipLaunchKernelGGL(k_census, dim3(ls.blocks[kStageCensus]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageCensus]);
    TE(kKCensus);
}

void launch_wta(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    size_t lds = (size_t)2 * (256 + ls.maxD - 1) * sizeof(uint64_t);
    TB(kKWta);
    hipLaunchKernelGGL(k_wta, dim3(ls.blocks[kStageWta]), dim3(256), lds, st, ls.rigs, ls.nrig, ls.bs[kStageWta]);
    TE(kKWta);
}

void launch_median(const LaunchSet& ls, hipStream_t st, KernelTimer* t)
{
    TB(kKMedianPrep);
    hipLaunchKernelGGL(k_median_prep, dim3(ls.blocks[kStageMedianPrep]), dim3(256), 0, st, ls.rigs, ls.nrig,
                       ls.bs[kStageMedianPrep], ls.scratch);
    TE(kKMedianPrep);
    size_t lds = (size_t)2 * ls.maxCols * sizeof(uint16_t);
    TB(kKMedianSerial);
    hipLaunchKernelGGL(k_median_serial, dim3(ls.nrig * 8), dim3(1024), lds, st, ls.rigs, ls.scratch);
    TE(kKMedianSerial);
}

void launch_mask(const LaunchSet& ls, int fmask, hipStream_t st, KernelTimer* t)
{
    TB(kKMask);
    hipLaunchKernelGGL(k_mask, dim3(ls.blocks[kStageMask]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageMask], fmask);
    TE(kKMask);
}

void launch_msum(const LaunchSet& ls, int fuse, hipStream_t st, KernelTimer* t)
{
    TB(kKMsum);
    hipLaunchKernelGGL(k_msum, dim3(ls.blocks[kStageMsum]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageMsum], fuse);
    TE(kKMsum);
}

void launch_sgm(const LaunchSet& ls, int mode, int keep_sum, hipStream_t st, KernelTimer* t)
{
    const int J = (ls.maxD + 63) / 64;
    TB(kKRows);
    SISTER_J_DISPATCH(J, k_sgm_rows, dim3(ls.blocks[kStageRows]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageRows]);
    TE(kKRows);
    TB(kKCols);
    SISTER_J_DISPATCH(J, k_sgm_cols, dim3(ls.blocks[kStageCols]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageCols]);
    TE(kKCols);
    TB(kKDiag);
    SISTER_J_DISPATCH(J, k_sgm_diag, dim3(ls.blocks[kStageDiag]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageDiag]);
    TE(kKDiag);
    TB(kKAnti);
    SISTER_J_DISPATCH(J, k_sgm_anti, dim3(ls.blocks[kStageDiag]), dim3(256), 0, st, ls.rigs, ls.nrig, ls.bs[kStageDiag],
                      mode, keep_sum);
    TE(kKAnti);
}

// per-rig block counts of each stage (the host builds the block-start tables)
void stage_blocks(const RigDesc& g, int out[kNumStages])
{
    const long npx = (long)g.Wp * g.Hp;
    out[kStageCensus] = 8 * ((g.Wp + 15) / 16) * ((g.Hp + 15) / 16);
    out[kStageWta] = 2 * g.Hp * ((g.Wp + 255) / 256) + 2 * g.Wp * ((g.Hp + 255) / 256);
    out[kStageMedianPrep] = (int)(8 * ((npx + 255) / 256));
    out[kStageMask] = (int)(4 * ((npx + 255) / 256));
    out[kStageMsum] = (int)((npx + 3) / 4);
    out[kStageRows] = (g.Hp + 3) / 4;
    out[kStageCols] = (g.Wp + 3) / 4;
    out[kStageDiag] = (g.Wp + g.Hp - 1 + 3) / 4;
}

}  // namespace sister
