/*
 * sister_oracle.c -- CPU restatement of SiSter's 5-view disparity path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product library (sister_amd/libsister_hip.so) never links or calls it.
 *
 * Every function restates one piece of the reference's CPU code, cited as
 * file:line relative to the reference checkout (cpp/...).  The restatement is
 * scalar C; it keeps the reference's observable semantics, including its
 * quirks:
 *   - census codes carry bit 0 of the previous raster pixel into bit 63
 *     (census.cpp:42-49, the accumulator is never reset);
 *   - cost rows R-2 and R-1 are never written by the reference
 *     (census.cpp:123-145); here they are 0, the value fresh zeroed pages give;
 *   - the 3x3 median runs in place, so each output reads the already
 *     filtered row above, and element C is forced to 0 (postprocess.cpp:15-71);
 *   - SGM truncates the first-line horizontal path to 8 bits via the
 *     uint8-argument saturate_cast (types.h:28, sgm.cpp:176,185,187), maps
 *     cost 255 to 0 only on first lines (sgm.cpp:109-110,146-147), restarts the
 *     horizontal path per row (sgm.cpp:215-216) and sees L=65535, minL=0
 *     beyond the left/right image border (sgm.cpp:57-81).
 *
 * Parity pins (see tests/test_oracle.py): golden vectors produced by the
 * reference's own kernels (census.cpp, sgm.cpp, postprocess.cpp compiled
 * unmodified into oracle/_ref), brute-force checks, and invariants.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#define SO_EXPORT __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* census.cpp:30-51 -- 9x7 center-symmetric census, rw=4, rh=3 (:154-155)    */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_census(const uint8_t* img, int rows, int cols, uint64_t* code)
{
    memset(code, 0, (size_t)rows * cols * sizeof(uint64_t));
    uint64_t acc = 0; /* never reset between pixels (census.cpp:42) */
    for (int i = 3; i < rows - 3; i++) {
        for (int j = 4; j < cols - 4; j++) {
            for (int y = -3; y <= 3; y++) {
                for (int x = -4; x <= 4; x++) {
                    int a = img[(size_t)(i + y) * cols + (j + x)];
                    int b = img[(size_t)(i - y) * cols + (j - x)];
                    acc = acc * 2u + (uint64_t)((a - b) > 0);
                }
            }
            code[(size_t)i * cols + j] = acc;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* census.cpp:54-146 -- Hamming cost volume, layout [row][col][d]            */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_cost(const uint64_t* cc, const uint64_t* cn, int rows, int cols,
                           int D, uint16_t* dsi)
{
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            uint16_t* cell = dsi + ((size_t)i * cols + j) * D;
            for (int d = 0; d < D; d++) {
                uint16_t v;
                if (i < 3) {
                    v = 255;                      /* :95-98 and :141-145 */
                } else if (i >= rows - 2) {
                    v = 0;                        /* never written (:123-135) */
                } else if (d > j) {
                    v = 255;                      /* :74-78 */
                } else {
                    uint64_t x = cc[(size_t)i * cols + j] ^ cn[(size_t)i * cols + (j - d)];
                    v = (uint16_t)__builtin_popcountll(x); /* :81-82 (4x popcount16 LUT) */
                }
                cell[d] = v;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* postprocess.cpp:74-185 -- WTALeft_SSE, uniqueness 1.0 (always accepted)   */
/* argmin over d <= min(j, D-1), lowest index on ties.                       */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_wta_left(const uint16_t* dsi, int rows, int cols, int D, float* out)
{
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            const uint16_t* c = dsi + ((size_t)i * cols + j) * D;
            int last = j < D - 1 ? j : D - 1;
            int best = 0;
            unsigned bestv = c[0];
            for (int d = 1; d <= last; d++)
                if (c[d] < bestv) { bestv = c[d]; best = d; }
            out[(size_t)i * cols + j] = (float)best;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* postprocess.cpp:187-315 -- WTARight_SSE: diagonal argmin                   */
/* argmin over k <= min(D-1, cols-1-j) of dsi[i][j+k][k], lowest index.      */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_wta_right(const uint16_t* dsi, int rows, int cols, int D, float* out)
{
    for (int i = 0; i < rows; i++) {
        for (int j = 0; j < cols; j++) {
            int last = (D - 1) < (cols - 1 - j) ? (D - 1) : (cols - 1 - j);
            int best = 0;
            unsigned bestv = dsi[((size_t)i * cols + j) * D];
            for (int k = 1; k <= last; k++) {
                unsigned v = dsi[((size_t)i * cols + j + k) * D + k];
                if (v < bestv) { bestv = v; best = k; }
            }
            out[(size_t)i * cols + j] = (float)best;
        }
    }
}

/* exact median of 9 (any exact method gives the value the sorting network */
/* of postprocess.cpp:52-58 returns)                                         */
static float med9(float* v)
{
    for (int a = 1; a < 9; a++) {
        float t = v[a];
        int b = a - 1;
        while (b >= 0 && v[b] > t) { v[b + 1] = v[b]; b--; }
        v[b + 1] = t;
    }
    return v[4];
}

/* ------------------------------------------------------------------------ */
/* postprocess.cpp:15-71 -- median3x3_SSE(source == dest), step for step.    */
/* The vector loop walks 4 flat positions at a time; each iteration reads   */
/* 8 floats from each of three lines starting at flat offset p, p+C, p+2C,   */
/* computes the 4 medians centred at p+C+1..p+C+4, and stores              */
/* [previous lane 3, lanes 0..2] at flat offset p+C (the alignr of :61).     */
/* Because source == dest, line 1 already holds filtered values.  The two   */
/* trailing memcpy calls (:69-70) copy the buffer onto itself.              */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_median_inplace(float* buf, int rows, int cols)
{
    const long C = cols;
    const long total = (long)rows * cols;
    float last3 = 0.0f;                    /* lastMedian = 0 (:29) */
    long p = 0;
    do {
        float med[4];
        for (int t = 0; t < 4; t++) {
            float v[9];
            for (int r = 0; r < 3; r++)
                for (int k = 0; k < 3; k++)
                    v[r * 3 + k] = buf[p + r * C + t + k];
            med[t] = med9(v);
        }
        buf[p + C + 0] = last3;
        buf[p + C + 1] = med[0];
        buf[p + C + 2] = med[1];
        buf[p + C + 3] = med[2];
        last3 = med[3];
        p += 4;
    } while (2 * C + p + 8 <= total);       /* line3+4+4 <= end (:67) */
}

/* ------------------------------------------------------------------------ */
/* postprocess.cpp:318-341 -- doLRCheck                                     */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_lrcheck(float* L, float* R, int rows, int cols, int thr)
{
    for (int i = 0; i < rows; i++) {
        float* l = L + (size_t)i * cols;
        const float* r = R + (size_t)i * cols;
        for (int j = 0; j < cols; j++) {
            float base = l[j];
            if (base >= 0 && base <= j) {
                float match = r[(int)(j - base)];
                int diff = (int)(base - match);
                if (abs(diff) > thr) l[j] = -10;
            } else {
                l[j] = -10;
            }
        }
    }
}

/* saturating uint16 helpers matching _mm_adds_epu16 / _mm_subs_epu16 */
static inline uint16_t adds16(unsigned a, unsigned b) { unsigned s = a + b; return (uint16_t)(s > 65535u ? 65535u : s); }
static inline uint16_t subs16(unsigned a, unsigned b) { return (uint16_t)(a > b ? a - b : 0u); }
static inline unsigned umin(unsigned a, unsigned b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* sgm.cpp:26-460 -- accumulateCostsSSE, 8 paths in 2 passes, P1=7,         */
/* P2 = adaptP2(alpha=0, gamma=100, P2min=17) = 100 (sgm.cpp:15-22,34-35).   */
/* State layout: per path a row of (W+2) pixel blocks of (D+2) uint16,       */
/* pixel -1 and W are border guards (all 65535, minL 0: sgm.cpp:57-81),      */
/* d slots -1 and D are 65535 (sgm.cpp:84-87,202-206).                      */
/* ------------------------------------------------------------------------ */
SO_EXPORT void oracle_sgm(const uint16_t* dsi, int H, int W, int D, uint16_t* S)
{
    const int P1 = 7, P2 = 100;
    const int st = D + 2;
    size_t rowlen = (size_t)(W + 2) * st;
    uint16_t* r0a = malloc(st * sizeof(uint16_t));
    uint16_t* r0b = malloc(st * sizeof(uint16_t));
    uint16_t* r1a = malloc(rowlen * sizeof(uint16_t));
    uint16_t* r1b = malloc(rowlen * sizeof(uint16_t));
    uint16_t* r2 = malloc(rowlen * sizeof(uint16_t));
    uint16_t* r3 = malloc(rowlen * sizeof(uint16_t));
    uint16_t* m1a = calloc(W + 2, sizeof(uint16_t));
    uint16_t* m1b = calloc(W + 2, sizeof(uint16_t));
    uint16_t* m2 = calloc(W + 2, sizeof(uint16_t));
    uint16_t* m3 = calloc(W + 2, sizeof(uint16_t));
    uint16_t* t0 = malloc(st * sizeof(uint16_t));
    uint16_t* t2 = malloc(st * sizeof(uint16_t));
    uint16_t* t3 = malloc(st * sizeof(uint16_t));
    for (int k = 0; k < st; k++) r0a[k] = r0b[k] = 65535;
    for (size_t k = 0; k < rowlen; k++) r1a[k] = r1b[k] = r2[k] = r3[k] = 65535;

#define BLK(buf, x) ((buf) + (size_t)((x) + 1) * st + 1) /* pointer to d=0 of pixel x */
#define CELL(base, y, x) ((base) + ((size_t)(y) * W + (x)) * D)

    for (int pass = 0; pass < 2; pass++) {
        const int dir = pass ? -1 : 1;
        const int y0 = pass ? H - 1 : 0;
        const int x0 = pass ? W - 1 : 0;
        uint16_t *r0p = r0a + 1, *r0c = r0b + 1;   /* L_r0_last / L_r0 */
        uint16_t *r1p = r1a, *r1c = r1b;           /* L_r1_last / L_r1 */
        uint16_t *m1p = m1a + 1, *m1c = m1b + 1;
        uint16_t* m2p = m2 + 1;
        uint16_t* m3p = m3 + 1;
        unsigned m0p;

        /* first line, first pixel (sgm.cpp:103-138) */
        {
            const uint16_t* c = CELL(dsi, y0, x0);
            uint16_t* s = CELL(S, y0, x0);
            unsigned mn = 65535;
            for (int d = 0; d < D; d++) {
                uint16_t v = c[d] == 255 ? 0 : c[d];
                r0p[d] = v;
                BLK(r1p, x0)[d] = v;
                BLK(r2, x0)[d] = v;
                BLK(r3, x0)[d] = v;
                if (v < mn) mn = v;
                s[d] = pass ? (uint16_t)(s[d] + v) : v;
            }
            m0p = mn;
            m1p[x0] = m2p[x0] = m3p[x0] = (uint16_t)mn;
        }
        /* rest of the first line (sgm.cpp:141-207) */
        for (int x = x0 + dir; x != (pass ? -1 : W); x += dir) {
            const uint16_t* c = CELL(dsi, y0, x);
            uint16_t* s = CELL(S, y0, x);
            unsigned mn = 65535, m0c = 65535;
            for (int d = 0; d < D; d++) {
                uint16_t v = c[d] == 255 ? 0 : c[d];
                BLK(r1p, x)[d] = v;
                BLK(r2, x)[d] = v;
                BLK(r3, x)[d] = v;
                if (v < mn) mn = v;
                int prop = r0p[d];
                int pm = (int)r0p[d - 1] + P1;
                int pp = (int)r0p[d + 1] + P1;
                if (prop > pm) prop = pm;
                if (prop > pp) prop = pp;
                if (prop > (int)m0p + P2) prop = (int)m0p + P2;
                prop -= (int)m0p;
                uint16_t nv = (uint8_t)(v + prop);   /* saturate_cast<uint16>(uint8) */
                r0c[d] = nv;
                if (nv < m0c) m0c = nv;
                s[d] = pass ? (uint16_t)(s[d] + nv) : nv;
            }
            m1p[x] = m2p[x] = m3p[x] = (uint16_t)mn;
            uint16_t* t = r0p; r0p = r0c; r0c = t;   /* swapPointers (:198-199) */
            m0p = m0c;
        }
        /* remaining lines (sgm.cpp:213-440) */
        for (int y = y0 + dir; y != (pass ? -1 : H); y += dir) {
            for (int d = 0; d < D; d++) r0p[d] = 0;   /* :215-216 */
            m0p = 0;
            for (int x = x0; x != (pass ? -1 : W); x += dir) {
                const uint16_t* c = CELL(dsi, y, x);
                uint16_t* s = CELL(S, y, x);
                const uint16_t* q1 = BLK(r1p, x - dir);   /* r1 from (y-dir, x-dir) */
                uint16_t* q2 = BLK(r2, x);                /* r2 from (y-dir, x), in place */
                uint16_t* q3 = BLK(r3, x + dir);          /* r3 from (y-dir, x+dir) */
                unsigned mm1 = m1p[x - dir], mm2 = m2p[x], mm3 = m3p[x + dir];
                unsigned p20 = adds16(P2, m0p), p21 = adds16(P2, mm1);
                unsigned p22 = adds16(P2, mm2), p23 = adds16(P2, mm3);
                unsigned n0 = 65535, n1 = 65535, n2 = 65535, n3 = 65535;
                uint16_t* o1 = BLK(r1c, x);
                for (int d = 0; d < D; d++) {
                    unsigned cv = c[d];
                    unsigned a0 = umin(umin(r0p[d], umin(adds16(r0p[d - 1], P1), adds16(r0p[d + 1], P1))), p20);
                    unsigned a1 = umin(umin(q1[d], umin(adds16(q1[d - 1], P1), adds16(q1[d + 1], P1))), p21);
                    unsigned a2 = umin(umin(q2[d], umin(adds16(q2[d - 1], P1), adds16(q2[d + 1], P1))), p22);
                    unsigned a3 = umin(umin(q3[d], umin(adds16(q3[d - 1], P1), adds16(q3[d + 1], P1))), p23);
                    uint16_t v0 = adds16(cv, subs16(a0, m0p));
                    uint16_t v1 = adds16(cv, subs16(a1, mm1));
                    uint16_t v2 = adds16(cv, subs16(a2, mm2));
                    uint16_t v3 = adds16(cv, subs16(a3, mm3));
                    t0[d] = v0; o1[d] = v1; t2[d] = v2; t3[d] = v3;
                    if (v0 < n0) n0 = v0;
                    if (v1 < n1) n1 = v1;
                    if (v2 < n2) n2 = v2;
                    if (v3 < n3) n3 = v3;
                    uint16_t tot = adds16(adds16(adds16(v0, v1), v2), v3);
                    s[d] = pass ? adds16(s[d], tot) : tot;
                }
                /* the SSE code stores each 8-wide block only after loading the
                   next one, so every d reads old values: store afterwards */
                for (int d = 0; d < D; d++) {
                    r0p[d] = t0[d];
                    q2[d] = t2[d];
                    BLK(r3, x)[d] = t3[d];
                }
                m0p = n0;
                m1c[x] = (uint16_t)n1;
                m2p[x] = (uint16_t)n2;
                m3p[x] = (uint16_t)n3;
            }
            { uint16_t* t = r1p; r1p = r1c; r1c = t; }   /* :437-438 */
            { uint16_t* t = m1p; m1p = m1c; m1c = t; }
        }
    }
#undef BLK
#undef CELL
    free(r0a); free(r0b); free(r1a); free(r1b); free(r2); free(r3);
    free(m1a); free(m1b); free(m2); free(m3); free(t0); free(t2); free(t3);
}

/* ------------------------------------------------------------------------ */
/* kernel table for the shared orchestration (orchestrate.c)                 */
/* ------------------------------------------------------------------------ */
#include "orchestrate.h"

/* ad_census (census.cpp:149-158): two census transforms, one cost volume */
static void k_cost(const uint8_t* imc, const uint8_t* imn, int rows, int cols, int D, uint16_t* dsi)
{
    size_t n = (size_t)rows * cols;
    uint64_t* cc = malloc(n * sizeof(uint64_t));
    uint64_t* cn = malloc(n * sizeof(uint64_t));
    oracle_census(imc, rows, cols, cc);
    oracle_census(imn, rows, cols, cn);
    oracle_cost(cc, cn, rows, cols, D, dsi);
    free(cc);
    free(cn);
}
static void k_wta_left(uint16_t* dsi, int r, int c, int D, float* o) { oracle_wta_left(dsi, r, c, D, o); }
static void k_wta_right(uint16_t* dsi, int r, int c, int D, float* o) { oracle_wta_right(dsi, r, c, D, o); }
static void k_sgm(uint16_t* dsi, const uint8_t* img, int H, int W, int D, uint16_t* S)
{
    (void)img; /* adaptP2 with alpha = 0 ignores the image (sgm.cpp:15-22) */
    oracle_sgm(dsi, H, W, D, S);
}

static const sister_cpu_kernels oracle_kernels = {
    k_cost, oracle_census, k_wta_left, k_wta_right,
    oracle_median_inplace, oracle_lrcheck, k_sgm,
};

SO_EXPORT int oracle_compute(const uint8_t* const views[5], int W, int H, int stride, int D,
                             int modes, uint16_t* out[3], float* subpix, uint16_t* disp,
                             sister_cpu_debug* dbg)
{
    return sister_cpu_compute(&oracle_kernels, views, W, H, stride, D, modes, out, subpix, disp, dbg);
}
