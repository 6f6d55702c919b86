/*
 * orchestrate.c -- CPU restatement of the helper class glue
 * (cpp/include/sister/SisterMultiviewDisparities.hpp:26-119, 152-295).
 *
 * TEST INFRASTRUCTURE ONLY (the oracle); see orchestrate.h.
 *
 * OpenCV semantics restated here (OpenCV is not available; parity of this
 * data movement against OpenCV itself is unpinned, it is plain indexing):
 *   copyMakeBorder(BORDER_REPLICATE)  -> clamp of coordinates (hpp:35-39)
 *   flip(.,1) / flip(transpose,-1) / flip(transpose,0)  (hpp:54-70)
 *   transpose + flip of the 90/270 masks back to original coords (hpp:235-236,250-251)
 *   convertTo(CV_16UC1) of integer-valued floats, crop Rect(D,D,W,H) and the
 *   saturating *255 (hpp:111-118).
 */
#include "orchestrate.h"
#include <stdlib.h>
#include <string.h>

static uint8_t* pad_replicate(const uint8_t* img, int W, int H, int stride, int D)
{
    int Wp = W + 2 * D, Hp = H + 2 * D;
    uint8_t* out = (uint8_t*)malloc((size_t)Wp * Hp);
    for (int y = 0; y < Hp; y++) {
        int sy = y - D; sy = sy < 0 ? 0 : (sy >= H ? H - 1 : sy);
        for (int x = 0; x < Wp; x++) {
            int sx = x - D; sx = sx < 0 ? 0 : (sx >= W ? W - 1 : sx);
            out[(size_t)y * Wp + x] = img[(size_t)sy * stride + sx];
        }
    }
    return out;
}

/* frame F image from a padded original (Wp x Hp).  Frame coords (r,c):
   F0: (y,x)=(r,c); F1 (90): flip(transpose,-1) -> (y,x)=(Hp-1-c, Wp-1-r);
   F2 (180): flip(.,1) -> (y,x)=(r, Wp-1-c); F3 (270): flip(transpose,0) ->
   (y,x)=(c, Wp-1-r). */
static uint8_t* make_frame(const uint8_t* pad, int Wp, int Hp, int F)
{
    int R = (F & 1) ? Wp : Hp, C = (F & 1) ? Hp : Wp;
    uint8_t* out = (uint8_t*)malloc((size_t)R * C);
    for (int r = 0; r < R; r++)
        for (int c = 0; c < C; c++) {
            int y, x;
            switch (F) {
            case 0: y = r; x = c; break;
            case 1: y = Hp - 1 - c; x = Wp - 1 - r; break;
            case 2: y = r; x = Wp - 1 - c; break;
            default: y = c; x = Wp - 1 - r; break;
            }
            out[(size_t)r * C + c] = pad[(size_t)y * Wp + x];
        }
    return out;
}

/* frame coords of original padded pixel (y,x) for frame F */
static inline size_t frame_index(int F, int y, int x, int Wp, int Hp)
{
    switch (F) {
    case 0: return (size_t)y * Wp + x;
    case 1: return (size_t)(Wp - 1 - x) * Hp + (Hp - 1 - y);
    case 2: return (size_t)y * Wp + (Wp - 1 - x);
    default: return (size_t)(Wp - 1 - x) * Hp + y;
    }
}

int sister_cpu_compute(const sister_cpu_kernels* k, const uint8_t* const views[5],
                       int W, int H, int stride, int D, int modes,
                       uint16_t* out[3], float* subpix, uint16_t* disp,
                       sister_cpu_debug* dbg)
{
    if (D < 8 || D % 8 != 0 || W % 4 != 0 || H % 4 != 0 || W < 4 || H < 4)
        return -1;
    const int Wp = W + 2 * D, Hp = H + 2 * D;
    const size_t npx = (size_t)Wp * Hp;
    const size_t ncell = npx * (size_t)D;

    /* hpp:35-39 pad; views order center,right,top,left,bottom; frame F pairs
       center with neighbour view index {1,2,3,4}[F] (right,top,left,bottom) */
    uint8_t* pad[5];
    for (int v = 0; v < 5; v++) pad[v] = pad_replicate(views[v], W, H, stride, D);
    uint8_t* fimg[8];
    for (int F = 0; F < 4; F++) {
        fimg[2 * F] = make_frame(pad[0], Wp, Hp, F);
        fimg[2 * F + 1] = make_frame(pad[1 + F], Wp, Hp, F);
    }

    /* hpp:181-184 -- four cost volumes, zero-filled (rows R-2,R-1 stay 0) */
    uint16_t* dsi[4];
    for (int F = 0; F < 4; F++) {
        int R = (F & 1) ? Wp : Hp, C = (F & 1) ? Hp : Wp;
        dsi[F] = (uint16_t*)calloc(ncell, sizeof(uint16_t));
        k->cost(fimg[2 * F], fimg[2 * F + 1], R, C, D, dsi[F]);
        if (dbg && k->census) {
            if (dbg->census[2 * F]) k->census(fimg[2 * F], R, C, dbg->census[2 * F]);
            if (dbg->census[2 * F + 1]) k->census(fimg[2 * F + 1], R, C, dbg->census[2 * F + 1]);
        }
    }

    /* hpp:193-252 -- confidence masks.  Identical in every mode that uses
       them, so they are computed once per needed frame. */
    int need[4] = {0, 0, 0, 0};
    if (modes & 3) need[0] = need[2] = 1;  /* mode != 2 */
    if (modes & 5) need[1] = need[3] = 1;  /* mode != 1 */
    uint8_t* mask[4];
    float* L = (float*)malloc(npx * sizeof(float));
    float* Rm = (float*)malloc(npx * sizeof(float));
    for (int F = 0; F < 4; F++) {
        mask[F] = (uint8_t*)calloc(npx, 1);
        if (!need[F]) continue;
        int R = (F & 1) ? Wp : Hp, C = (F & 1) ? Hp : Wp;
        k->wta_left(dsi[F], R, C, D, L);
        k->wta_right(dsi[F], R, C, D, Rm);
        if (dbg && dbg->wta_l[F]) memcpy(dbg->wta_l[F], L, npx * sizeof(float));
        if (dbg && dbg->wta_r[F]) memcpy(dbg->wta_r[F], Rm, npx * sizeof(float));
        k->median_inplace(L, R, C);
        k->median_inplace(Rm, R, C);
        if (dbg && dbg->med_l[F]) memcpy(dbg->med_l[F], L, npx * sizeof(float));
        if (dbg && dbg->med_r[F]) memcpy(dbg->med_r[F], Rm, npx * sizeof(float));
        k->lrcheck(L, Rm, R, C, 5);
        for (int y = 0; y < Hp; y++)
            for (int x = 0; x < Wp; x++) {
                size_t q = frame_index(F, y, x, Wp, Hp);
                int c = (int)(q % (size_t)C);
                mask[F][(size_t)y * Wp + x] = (L[q] <= 0 || c < D) ? 0 : 1;  /* hpp:203 */
            }
        if (dbg && dbg->mask[F]) memcpy(dbg->mask[F], mask[F], npx);
    }
    free(L); free(Rm);

    uint16_t* msum = (uint16_t*)malloc(ncell * sizeof(uint16_t));
    uint16_t* aligned[4] = {NULL, NULL, NULL, NULL};
    for (int F = 1; F < 4; F++) aligned[F] = (uint16_t*)malloc(ncell * sizeof(uint16_t));
    uint16_t* S = (uint16_t*)calloc(ncell, sizeof(uint16_t));
    float* dp = (float*)malloc(npx * sizeof(float));
    for (int mode = 0; mode < 3; mode++) {
        if (!(modes & (1 << mode))) continue;
        /* hpp:255-277 -- realign the 180/90/270 volumes into original
           coordinates (three full copies, as the reference writes them) and
           take the mask-weighted sum */
        int use[4];
        use[0] = use[2] = (mode != 2);
        use[1] = use[3] = (mode != 1);
        for (int y = 0; y < Hp; y++)
            for (int x = 0; x < Wp; x++) {
                size_t cell = ((size_t)y * Wp + x) * D;
                uint16_t* o = msum + cell;
                const uint16_t* src[4];
                int m[4];
                for (int F = 0; F < 4; F++) {
                    src[F] = dsi[F] + frame_index(F, y, x, Wp, Hp) * D;
                    m[F] = use[F] ? mask[F][(size_t)y * Wp + x] : 0;
                }
                for (int d = 0; d < D; d++) {
                    for (int F = 1; F < 4; F++) aligned[F][cell + d] = src[F][d];
                    int acc = m[0] * src[0][d];
                    for (int F = 1; F < 4; F++) acc += m[F] * aligned[F][cell + d];
                    o[d] = (uint16_t)acc;
                }
            }
        if (dbg && dbg->msum[mode]) memcpy(dbg->msum[mode], msum, ncell * sizeof(uint16_t));
        /* hpp:280 -- SGM; hpp:283 -- WTA on the aggregate */
        k->sgm(msum, pad[0], Hp, Wp, D, S);
        if (dbg && dbg->sum[mode]) memcpy(dbg->sum[mode], S, ncell * sizeof(uint16_t));
        k->wta_left(S, Hp, Wp, D, dp);
        if (dbg && dbg->disp_pad[mode]) memcpy(dbg->disp_pad[mode], dp, npx * sizeof(float));
        /* hpp:111-118 -- convertTo(CV_16U), crop, *255 with saturation */
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                float f = dp[(size_t)(y + D) * Wp + (x + D)];
                unsigned v = (unsigned)f;
                unsigned s = v * 255u;
                if (out && out[mode]) out[mode][(size_t)y * W + x] = (uint16_t)(s > 65535u ? 65535u : s);
                if (mode == 0 && disp) disp[(size_t)y * W + x] = (uint16_t)v;
                if (mode == 0 && subpix) {
                    /* build-defined parabola refinement (absent in the reference) */
                    const uint16_t* c = S + ((size_t)(y + D) * Wp + (x + D)) * D;
                    double delta = 0.0;
                    if (v > 0 && (int)v < D - 1) {
                        double sm = c[v - 1], s0 = c[v], sp = c[v + 1];
                        double den = sm - 2.0 * s0 + sp;
                        if (den > 0) delta = (sm - sp) / (2.0 * den);
                    }
                    subpix[(size_t)y * W + x] = (float)((double)v + delta);
                }
            }
    }
    free(msum); free(S); free(dp);
    for (int F = 1; F < 4; F++) free(aligned[F]);
    for (int F = 0; F < 4; F++) { free(dsi[F]); free(mask[F]); }
    for (int i = 0; i < 8; i++) free(fimg[i]);
    for (int v = 0; v < 5; v++) free(pad[v]);
    return 0;
}
