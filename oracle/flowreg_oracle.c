/*
 * flowreg_oracle.c -- CPU restatement of pyflowreg's variational optical-flow
 * path (get_displacement with the gradient-constancy data term).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be built and
 * called only by tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py.  The product (pyflowreg_amd) never links or calls it.
 *
 * Compiled twice by oracle/Makefile:
 *   liboracle64.so  -- OR_REAL=double : the FP64 restatement (parity oracle)
 *   liboracle32.so  -- OR_REAL=float  : the same arithmetic in FP32, used only
 *                      for the FP32-vs-FP64 precision study.
 *
 * Float32 islands follow the reference exactly in both builds: level images,
 * warped images, resized flows and resized weights are float32-valued; the
 * resize accumulates float32 products in float64.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the pyflowreg repository root).  Arrays are C-contiguous, interleaved
 * (rows, cols, channels) like the reference's numpy arrays.
 *
 * Pinning (what checks each function; tests/test_oracle.py unless named):
 *   or_warping_depth        closed form + level tables (19 levels at 512^2,
 *                           16 / 21 / 22 at 256^2, 796x512, 1024^2)
 *   or_fill_resize_tables   partition of unity, indices in range
 *   or_resize_tab           identity at scale 1; upsampling = torch's Keys
 *                           weights on the reference's 3-tap window;
 *                           downsampling = scipy-style float32 Gaussian then
 *                           torch's Keys resampling, minus the dropped tap
 *   or_warp_cubic(64)       reference warp tests (zero / integer shifts,
 *                           out-of-bounds <- fixed); torch bicubic at the
 *                           1/32-px-rounded positions (a 0.3 px shift lands
 *                           at 0.3125); the 64F sums against numpy float64
 *   or_warp_linear(64)      the same reference tests; torch bilinear at the
 *                           rounded positions (tests/test_xcorr_oracle.py)
 *   16-bit frames           (oracle.py imregister_wrapper over or_warp_cubic /
 *                           or_warp_linear) identity at the type's extremes,
 *                           half-pixel rounding half to even, clamped cubic
 *                           overshoot
 *   or_motion_tensor_gc     reference known answer J11=0.8, identical frames
 *                           give J33=J13=J23=0, quadratic closed forms,
 *                           numpy's float32 arithmetic for the warped image
 *   or_compute_flow         f1=f2 gives zero flow; a_data=1 converges to the
 *                           dense direct solve; 1-3 sweeps equal the matrix
 *                           form of SOR in column-major order; the literal
 *                           per-sweep expression order of the reference
 *   or_cubic_resize_cv      torch's Keys (A = -0.75) weights at OpenCV's
 *                           float32 source coordinates (512^2, 256^2, 796x512
 *                           from min_level 4 and 6), constants, identity at
 *                           scale 1.  Which table cell OpenCV's own float
 *                           coefficient evaluation rounds differently is
 *                           parity unpinned (cv2 absent).
 *   (the level loop, the median and the Gaussian prefilter live in
 *   oracle/oracle.py; the median and Gaussian call scipy, the reference's
 *   own dependency, directly)
 */
#include <tgmath.h>
#include <stdint.h>
#include <stdlib.h>
#include <stddef.h>
#include <string.h>

#include "../include/flowreg_detmath.h"

#ifndef OR_REAL
#define OR_REAL double
#endif
typedef OR_REAL real;

#define KEYS_A (-0.75)

/* Python's round(): half to even (default FP rounding mode). */
static double py_round(double x) { return nearbyint(x); }

/* core/warping.py:237-277 (warpingDepth) */
int or_warping_depth(double eta, int levels, int m, int n)
{
    double min_dim = (double)(m < n ? m : n);
    int depth = 0;
    for (int k = 0; k < levels; ++k) {
        depth += 1;
        min_dim *= eta;
        if (py_round(min_dim) < 10.0) break;
    }
    return depth;
}

/* util/resize_util.py:68-76 (_cubic, A=-0.75) */
static double keys_cubic(double x)
{
    double ax = fabs(x);
    if (ax < 1.0) return (KEYS_A + 2.0) * ax * ax * ax - (KEYS_A + 3.0) * ax * ax + 1.0;
    if (ax < 2.0) return KEYS_A * ax * ax * ax - 5.0 * KEYS_A * ax * ax + 8.0 * KEYS_A * ax - 4.0 * KEYS_A;
    return 0.0;
}

/* util/resize_util.py:182-191 (_reflect_idx: ...b a | a b...) */
static int reflect_idx(int j, int n)
{
    if (n <= 1) return 0;
    while (j < 0 || j >= n) {
        if (j < 0) j = -j - 1;
        else j = 2 * n - 1 - j;
    }
    return j;
}

/* util/resize_util.py:216-229 (_precompute_fused_gauss_cubic): tap radius. */
int or_resize_radius(double sigma)
{
    if (sigma <= 0.0) return 0;
    return (int)ceil(2.0 * sigma);
}

/* util/resize_util.py:194-213 (_fill_tables_fused_gauss_cubic_reflect):
 * fused Gaussian+cubic tap tables, idx/wt are (out_len, 2R+4).  g is the
 * float32 Gaussian of util/resize_util.py:216-225, computed by the caller
 * with numpy exactly as the reference does. */
void or_fill_resize_tables(int in_len, int out_len, int R, const float *g, int32_t *idx, float *wt)
{
    int P = 2 * R + 4;
    double scale = (double)out_len / (double)in_len;
    for (int i = 0; i < out_len; ++i) {
        double x = (i + 0.5) / scale - 0.5;
        int left = (int)floor(x - 2.0) - R;
        double ssum = 0.0;
        for (int p = 0; p < P; ++p) {
            int j = left + p;
            idx[(size_t)i * P + p] = reflect_idx(j, in_len);
            double d = x - j, acc = 0.0;
            for (int u = -R; u <= R; ++u) acc += (double)g[u + R] * keys_cubic(d - u);
            wt[(size_t)i * P + p] = (float)acc;
            ssum += acc;
        }
        double inv = 1.0 / ssum;
        for (int p = 0; p < P; ++p) wt[(size_t)i * P + p] = (float)((double)wt[(size_t)i * P + p] * inv);
    }
}

/* util/resize_util.py:116-141 (_resize_h/_resize_v) + :232-254
 * (imresize_fused_gauss_cubic), one channel of an interleaved (h,w,C)
 * float32 image.  Products are float32, sums float64, stores float32. */
static void resize_channel(const float *src, int h, int w, int C, int c,
                           float *dst, int oh, int ow,
                           const int32_t *ix, const float *wx, int Px,
                           const int32_t *iy, const float *wy, int Py, float *tmp)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < ow; ++x) {
            double s = 0.0;
            for (int p = 0; p < Px; ++p) {
                float prod = src[((size_t)y * w + ix[(size_t)x * Px + p]) * C + c] * wx[(size_t)x * Px + p];
                s += (double)prod;
            }
            tmp[(size_t)y * ow + x] = (float)s;
        }
    for (int x = 0; x < ow; ++x)
        for (int y = 0; y < oh; ++y) {
            double s = 0.0;
            for (int p = 0; p < Py; ++p) {
                float prod = tmp[(size_t)iy[(size_t)y * Py + p] * ow + x] * wy[(size_t)y * Py + p];
                s += (double)prod;
            }
            dst[((size_t)y * ow + x) * C + c] = (float)s;
        }
}

/* util/resize_util.py:243-254: separable resize of an interleaved (h,w,C)
 * float32 image with precomputed tables (P taps per output sample). */
void or_resize_tab(const float *src, int h, int w, int C, float *dst, int oh, int ow,
                   const int32_t *ix, const float *wx, const int32_t *iy, const float *wy, int P)
{
    float *tmp = (float *)malloc(sizeof(float) * (size_t)h * ow);
    for (int c = 0; c < C; ++c) resize_channel(src, h, w, C, c, dst, oh, ow, ix, wx, P, iy, wy, P, tmp);
    free(tmp);
}

/* OpenCV's float-table bicubic coefficients (Keys, A=-0.75) at t = k/32,
 * evaluated in float as OpenCV's interpolateCubic does. */
static void cv_cubic_coeffs(float x, float *c)
{
    const float A = -0.75f;
    c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
    c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
    c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
    c[3] = 1.f - c[0] - c[1] - c[2];
}

/* core/warping.py:160-234 (imregister_wrapper, cubic):
 *   map = float32(grid + u); OOB on the unclipped map; clip; cv2.remap with
 *   INTER_CUBIC + BORDER_REPLICATE into float32; OOB pixels <- f1.
 * cv2.remap (opencv-contrib-python>=4.10, third party, restated from its
 * public behaviour; parity with OpenCV itself is unpinned): float maps are
 * rounded to 1/32 px (cvRound, half-to-even), the 4x4 weights are products
 * of two float 1-D tables.  The sums follow the image depth: float for
 * float32 images (or_warp_cubic, OpenCV's 32F path), double for float64
 * images (or_warp_cubic64, its 64F path: float weights, double products and
 * sums); get_displacement's level images are float64, so its level warps
 * take the 64F path.  Interior pixels sum row by row, border pixels
 * (BORDER_REPLICATE) sum the 16 taps sequentially.  The reference stores the
 * result, and the OOB fill, in a float32 array.
 * u, v are the per-level displacements already divided by the grid spacing
 * (u / h_col, v / h_row in the reference), as float64. */
static int g_warp_quantize = 1;
/* Study switch: 0 evaluates the bicubic weights at the exact fractional
 * position instead of OpenCV's 1/32-px grid (never used for parity). */
void or_set_warp_quantize(int q) { g_warp_quantize = q; }

/* Map of pixel (x, y): 0 when out of bounds, else 1 with the tap origin
 * (sx, sy), the 16 weights and whether all taps are inside (interior). */
static int warp_taps(int x, int y, int H, int W, double u, double v, float tab[32][4], int *sx, int *sy,
                     float w[16], int *interior)
{
    float mx = (float)((double)x + u);
    float my = (float)((double)y + v);
    if ((mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H)) return 0;
    float cx = mx < 0.f ? 0.f : (mx > (float)(W - 1) ? (float)(W - 1) : mx);
    float cy = my < 0.f ? 0.f : (my > (float)(H - 1) ? (float)(H - 1) : my);
    int X = (int)lrintf(cx * 32.f), Y = (int)lrintf(cy * 32.f);
    *sx = (X >> 5) - 1;
    *sy = (Y >> 5) - 1;
    const float *wxv = tab[X & 31], *wyv = tab[Y & 31];
    float cxq[4], cyq[4];
    if (!g_warp_quantize) {
        *sx = (int)floorf(cx) - 1; *sy = (int)floorf(cy) - 1;
        cv_cubic_coeffs(cx - floorf(cx), cxq); cv_cubic_coeffs(cy - floorf(cy), cyq);
        wxv = cxq; wyv = cyq;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) w[i * 4 + j] = wyv[i] * wxv[j];
    int width1 = W - 3 > 0 ? W - 3 : 0, height1 = H - 3 > 0 ? H - 3 : 0;
    *interior = (unsigned)*sx < (unsigned)width1 && (unsigned)*sy < (unsigned)height1;
    return 1;
}

static void border_taps(int sx, int sy, int H, int W, int xs[4], int ys[4])
{
    for (int i = 0; i < 4; ++i) {
        int a = sx + i, b = sy + i;
        xs[i] = a < 0 ? 0 : (a >= W ? W - 1 : a);
        ys[i] = b < 0 ? 0 : (b >= H ? H - 1 : b);
    }
}

void or_warp_cubic(const float *f2, const float *f1, int H, int W, int C,
                   const double *u, const double *v, float *out)
{
    float tab[32][4];
    for (int k = 0; k < 32; ++k) cv_cubic_coeffs(k * (1.f / 32), tab[k]);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            int sx, sy, interior;
            float w[16];
            if (!warp_taps(x, y, H, W, u[p], v[p], tab, &sx, &sy, w, &interior)) {
                for (int c = 0; c < C; ++c) out[p * C + c] = f1[p * C + c];
                continue;
            }
            for (int c = 0; c < C; ++c) {
                float sum;
                if (interior) {
                    const float *S = f2 + ((size_t)sy * W + sx) * C + c;
                    size_t rs = (size_t)W * C;
                    sum = S[0] * w[0] + S[C] * w[1] + S[2 * C] * w[2] + S[3 * C] * w[3];
                    S += rs;
                    sum += S[0] * w[4] + S[C] * w[5] + S[2 * C] * w[6] + S[3 * C] * w[7];
                    S += rs;
                    sum += S[0] * w[8] + S[C] * w[9] + S[2 * C] * w[10] + S[3 * C] * w[11];
                    S += rs;
                    sum += S[0] * w[12] + S[C] * w[13] + S[2 * C] * w[14] + S[3 * C] * w[15];
                } else {
                    int xs[4], ys[4];
                    border_taps(sx, sy, H, W, xs, ys);
                    sum = 0.f;
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j)
                            sum += f2[((size_t)ys[i] * W + xs[j]) * C + c] * w[i * 4 + j];
                }
                out[p * C + c] = sum;
            }
        }
}

void or_warp_cubic64(const double *f2, const double *f1, int H, int W, int C,
                     const double *u, const double *v, float *out)
{
    float tab[32][4];
    for (int k = 0; k < 32; ++k) cv_cubic_coeffs(k * (1.f / 32), tab[k]);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            int sx, sy, interior;
            float w[16];
            if (!warp_taps(x, y, H, W, u[p], v[p], tab, &sx, &sy, w, &interior)) {
                for (int c = 0; c < C; ++c) out[p * C + c] = (float)f1[p * C + c];
                continue;
            }
            for (int c = 0; c < C; ++c) {
                double sum;
                if (interior) {
                    const double *S = f2 + ((size_t)sy * W + sx) * C + c;
                    size_t rs = (size_t)W * C;
                    sum = S[0] * (double)w[0] + S[C] * (double)w[1] + S[2 * C] * (double)w[2] +
                          S[3 * C] * (double)w[3];
                    S += rs;
                    sum += S[0] * (double)w[4] + S[C] * (double)w[5] + S[2 * C] * (double)w[6] +
                           S[3 * C] * (double)w[7];
                    S += rs;
                    sum += S[0] * (double)w[8] + S[C] * (double)w[9] + S[2 * C] * (double)w[10] +
                           S[3 * C] * (double)w[11];
                    S += rs;
                    sum += S[0] * (double)w[12] + S[C] * (double)w[13] + S[2 * C] * (double)w[14] +
                           S[3 * C] * (double)w[15];
                } else {
                    int xs[4], ys[4];
                    border_taps(sx, sy, H, W, xs, ys);
                    sum = 0.0;
                    for (int i = 0; i < 4; ++i)
                        for (int j = 0; j < 4; ++j)
                            sum += f2[((size_t)ys[i] * W + xs[j]) * C + c] * (double)w[i * 4 + j];
                }
                out[p * C + c] = (float)sum;
            }
        }
}

/* cv2.remap on 8-bit images (opencv-contrib-python>=4.10, third party;
 * imgwarp.cpp initInterTab2D / remapBicubic / remapBilinear with
 * FixedPtCast<int, uchar, 15>, restated; parity with OpenCV itself is
 * unpinned).  The 2-D weights of table cell (fx, fy) are the float products
 * wy[k1] * wx[k2] of the 1-D float table, scaled by 2^15 and rounded to
 * int16 (cvRound, saturated: a weight of exactly 1 becomes 32767).  When
 * they do not sum to 2^15, the difference is taken off the largest (sum too
 * small) or put on the smallest (sum too large) of the cells
 * (k1, k2) in [ks/2, ks/2 + 2)^2, scanned row by row, first extreme wins.
 * For the bilinear 2x2 table that window reaches past the block; OpenCV
 * reads the next, still unfilled (zero) table cell there, restated as
 * zeros.  The sample is the exact integer sum of the taps (clamped
 * indices, BORDER_REPLICATE), stored as saturate((sum + 2^14) >> 15). */
void or_remap_u8_weights(int linear, int fx, int fy, int *itab)
{
    const int ks = linear ? 2 : 4;
    float tx[4], ty[4];
    if (linear) {
        tx[1] = (float)fx * (1.f / 32); tx[0] = 1.f - tx[1];
        ty[1] = (float)fy * (1.f / 32); ty[0] = 1.f - ty[1];
    } else {
        cv_cubic_coeffs((float)fx * (1.f / 32), tx);
        cv_cubic_coeffs((float)fy * (1.f / 32), ty);
    }
    int buf[24] = {0}, isum = 0;
    for (int k1 = 0; k1 < ks; ++k1)
        for (int k2 = 0; k2 < ks; ++k2) {
            float wv = ty[k1] * tx[k2];
            long r = lrintf(wv * 32768.f);
            r = r < -32768 ? -32768 : (r > 32767 ? 32767 : r);
            buf[k1 * ks + k2] = (int)r;
            isum += (int)r;
        }
    if (isum != 32768) {
        int diff = isum - 32768, h = ks / 2, M1 = h, M2 = h, m1 = h, m2 = h;
        for (int k1 = h; k1 < h + 2; ++k1)
            for (int k2 = h; k2 < h + 2; ++k2) {
                if (buf[k1 * ks + k2] < buf[m1 * ks + m2]) m1 = k1, m2 = k2;
                else if (buf[k1 * ks + k2] > buf[M1 * ks + M2]) M1 = k1, M2 = k2;
            }
        if (diff < 0) buf[M1 * ks + M2] = (int16_t)(buf[M1 * ks + M2] - diff);
        else buf[m1 * ks + m2] = (int16_t)(buf[m1 * ks + m2] - diff);
    }
    for (int k = 0; k < ks * ks; ++k) itab[k] = buf[k];
}

void or_warp_u8(int linear, const uint8_t *f2, const float *f1, int H, int W, int C,
                const double *u, const double *v, float *out)
{
    const int ks = linear ? 2 : 4, off = linear ? 0 : 1;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            float mx = (float)((double)x + u[p]);
            float my = (float)((double)y + v[p]);
            if ((mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H)) {
                for (int c = 0; c < C; ++c) out[p * C + c] = f1[p * C + c];
                continue;
            }
            float cx = mx > (float)(W - 1) ? (float)(W - 1) : mx;
            float cy = my > (float)(H - 1) ? (float)(H - 1) : my;
            int X = (int)lrintf(cx * 32.f), Y = (int)lrintf(cy * 32.f);
            int sx = (X >> 5) - off, sy = (Y >> 5) - off, itab[16], xs[4], ys[4];
            or_remap_u8_weights(linear, X & 31, Y & 31, itab);
            for (int i = 0; i < ks; ++i) {
                int a = sx + i, b = sy + i;
                xs[i] = a < 0 ? 0 : (a >= W ? W - 1 : a);
                ys[i] = b < 0 ? 0 : (b >= H ? H - 1 : b);
            }
            for (int c = 0; c < C; ++c) {
                int sum = 0;
                for (int i = 0; i < ks; ++i)
                    for (int j = 0; j < ks; ++j)
                        sum += (int)f2[((size_t)ys[i] * W + xs[j]) * C + c] * itab[i * ks + j];
                int r = (sum + (1 << 14)) >> 15;
                out[p * C + c] = (float)(r < 0 ? 0 : (r > 255 ? 255 : r));
            }
        }
}

/* core/warping.py:160-234 with interpolation_method="linear": cv2.remap
 * INTER_LINEAR + BORDER_REPLICATE, restated like the cubic path above
 * (parity with OpenCV itself unpinned).  The map is rounded to 1/32 px; the
 * tap origin is (X >> 5, Y >> 5) and the 2x2 weights are products
 * wy[i] * wx[j] of the float table {1 - k/32, k/32} (all exact).  A pixel
 * whose origin satisfies sx < W-1 and sy < H-1 sums its four taps left to
 * right; elsewhere the taps are clamped to the image (BORDER_REPLICATE) and
 * summed in the same order.  Float sums for float32 images, double sums of
 * double products for float64 images. */
static int lin_taps(int x, int y, int H, int W, double u, double v, int xs[2], int ys[2], float w[4])
{
    float mx = (float)((double)x + u);
    float my = (float)((double)y + v);
    if ((mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H)) return 0;
    float cx = mx > (float)(W - 1) ? (float)(W - 1) : mx;
    float cy = my > (float)(H - 1) ? (float)(H - 1) : my;
    int X = (int)lrintf(cx * 32.f), Y = (int)lrintf(cy * 32.f);
    int sx = X >> 5, sy = Y >> 5;
    float fx = (float)(X & 31) * (1.f / 32), fy = (float)(Y & 31) * (1.f / 32);
    float wx[2] = {1.f - fx, fx}, wy[2] = {1.f - fy, fy};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) w[i * 2 + j] = wy[i] * wx[j];
    for (int i = 0; i < 2; ++i) {
        int a = sx + i, b = sy + i;
        xs[i] = a < 0 ? 0 : (a >= W ? W - 1 : a);
        ys[i] = b < 0 ? 0 : (b >= H ? H - 1 : b);
    }
    return 1;
}

void or_warp_linear(const float *f2, const float *f1, int H, int W, int C,
                    const double *u, const double *v, float *out)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            int xs[2], ys[2];
            float w[4];
            if (!lin_taps(x, y, H, W, u[p], v[p], xs, ys, w)) {
                for (int c = 0; c < C; ++c) out[p * C + c] = f1[p * C + c];
                continue;
            }
            for (int c = 0; c < C; ++c) {
                const float *S = f2 + c;
                float sum = S[((size_t)ys[0] * W + xs[0]) * C] * w[0] + S[((size_t)ys[0] * W + xs[1]) * C] * w[1] +
                            S[((size_t)ys[1] * W + xs[0]) * C] * w[2] + S[((size_t)ys[1] * W + xs[1]) * C] * w[3];
                out[p * C + c] = sum;
            }
        }
}

void or_warp_linear64(const double *f2, const double *f1, int H, int W, int C,
                      const double *u, const double *v, float *out)
{
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t p = (size_t)y * W + x;
            int xs[2], ys[2];
            float w[4];
            if (!lin_taps(x, y, H, W, u[p], v[p], xs, ys, w)) {
                for (int c = 0; c < C; ++c) out[p * C + c] = (float)f1[p * C + c];
                continue;
            }
            for (int c = 0; c < C; ++c) {
                const double *S = f2 + c;
                double sum = S[((size_t)ys[0] * W + xs[0]) * C] * (double)w[0] +
                             S[((size_t)ys[0] * W + xs[1]) * C] * (double)w[1] +
                             S[((size_t)ys[1] * W + xs[0]) * C] * (double)w[2] +
                             S[((size_t)ys[1] * W + xs[1]) * C] * (double)w[3];
                out[p * C + c] = (float)sum;
            }
        }
}

/* ---------------------------------------------------------------------
 * Motion tensors (core/optical_flow.py:73-345).  numpy evaluates each
 * expression in the dtype of its array operands: the level image f1 is
 * float64 (the resize casts back to the input dtype), the warped image f2
 * is float32 below the top level (imregister_wrapper returns float32) and
 * float64 at the top level (a copy of the resized level).  p1 / p2 flag a
 * float32 input; an operation on float32 operands only is rounded to
 * float32, and one with a float64 operand is float64.  Rounding the double
 * result of one +, -, *, / or sqrt of float32 values to float32 equals the
 * float32 operation (double has more than 2 * 24 + 2 bits).  Python float
 * scalars are cast to the array dtype first (numpy's weak scalars).
 * --------------------------------------------------------------------- */
static double r32(int p, double x) { return p ? (double)(float)x : x; }

/* np.gradient interior value (x[+1] - x[-1]) / (2. * h) */
static double t_grad(int p, double xp, double xm, double h)
{
    return r32(p, r32(p, xp - xm) / r32(p, 2.0 * h));
}

/* (x[-1] - 2 * x[0] + x[+1]) / (h ** 2) */
static double t_second(int p, double xm, double x0, double xp, double h)
{
    return r32(p, r32(p, r32(p, xm - 2 * x0) + xp) / r32(p, h * h));
}

/* symmetric / edge pad of an (h, w) channel: the padded ring copies its
 * neighbour (the same for mode="symmetric" and "edge" with width 1) */
static double *t_pad(const double *f, int h, int w, int C, int c)
{
    int hp = h + 2, wp = w + 2;
    double *F = (double *)malloc(sizeof(double) * (size_t)hp * wp);
    for (int j = 0; j < hp; ++j)
        for (int i = 0; i < wp; ++i) {
            int sj = j - 1 < 0 ? 0 : (j - 1 >= h ? h - 1 : j - 1);
            int si = i - 1 < 0 ? 0 : (i - 1 >= w ? w - 1 : i - 1);
            F[(size_t)j * wp + i] = f[((size_t)sj * w + si) * C + c];
        }
    return F;
}

static void t_zero_ring(int hp, int wp, real *J11, real *J22, real *J33, real *J12, real *J13,
                        real *J23)
{
    for (int j = 0; j < hp; ++j)
        for (int i = 0; i < wp; ++i)
            if (j == 0 || j == hp - 1 || i == 0 || i == wp - 1) {
                size_t q = (size_t)j * wp + i;
                J11[q] = J22[q] = J33[q] = J12[q] = J13[q] = J23[q] = 0;
            }
}

/* core/optical_flow.py:73-160 (get_motion_tensor_gc) for one channel.
 * f1, f2: (h, w) images (strided by C); p1 / p2 mark a float32 dtype.
 * Outputs J* (h+2, w+2) with a zero border. */
void or_motion_tensor_gc(const double *f1, const double *f2, int h, int w, int C, int c,
                         double hy, double hx, int p1, int p2,
                         real *J11, real *J22, real *J33, real *J12, real *J13, real *J23)
{
    const int pm = p1 && p2;
    int hp = h + 2, wp = w + 2;
    size_t np_ = (size_t)hp * wp;
    double *F1 = t_pad(f1, h, w, C, c), *F2 = t_pad(f2, h, w, C, c);
    double *fx = (double *)malloc(sizeof(double) * np_), *ft = (double *)malloc(sizeof(double) * np_);
    /* fx = 0.5 (d/dx F1 + d/dx F2), ft = F2 - F1 on the interior; the ring
     * is replaced by the symmetric re-pad */
    for (int j = 1; j <= h; ++j)
        for (int i = 1; i <= w; ++i) {
            size_t q = (size_t)j * wp + i;
            double g1 = t_grad(p1, F1[q + 1], F1[q - 1], hx);
            double g2 = t_grad(p2, F2[q + 1], F2[q - 1], hx);
            fx[q] = r32(pm, 0.5 * r32(pm, g1 + g2));
            ft[q] = r32(pm, F2[q] - F1[q]);
        }
    for (int j = 0; j < hp; ++j)
        for (int i = 0; i < wp; ++i) {
            if (j >= 1 && j <= h && i >= 1 && i <= w) continue;
            int sj = j < 1 ? 1 : (j > h ? h : j);
            int si = i < 1 ? 1 : (i > w ? w : i);
            fx[(size_t)j * wp + i] = fx[(size_t)sj * wp + si];
            ft[(size_t)j * wp + i] = ft[(size_t)sj * wp + si];
        }
    for (int j = 1; j <= h; ++j)
        for (int i = 1; i <= w; ++i) {
            size_t q = (size_t)j * wp + i;
            double fxy = t_grad(pm, fx[q + wp], fx[q - wp], hy);
            double fxt = t_grad(pm, ft[q + 1], ft[q - 1], hx);
            double fyt = t_grad(pm, ft[q + wp], ft[q - wp], hy);
            double fxx1 = t_second(p1, F1[q - 1], F1[q], F1[q + 1], hx);
            double fyy1 = t_second(p1, F1[q - wp], F1[q], F1[q + wp], hy);
            double fxx2 = t_second(p2, F2[q - 1], F2[q], F2[q + 1], hx);
            double fyy2 = t_second(p2, F2[q - wp], F2[q], F2[q + wp], hy);
            double fxx = r32(pm, 0.5 * r32(pm, fxx1 + fxx2));
            double fyy = r32(pm, 0.5 * r32(pm, fyy1 + fyy2));
            /* reg = 1.0 / ((np.sqrt(a**2 + b**2) ** 2) + 1e-6) */
            double sx = r32(pm, sqrt(r32(pm, r32(pm, fxx * fxx) + r32(pm, fxy * fxy))));
            double sy = r32(pm, sqrt(r32(pm, r32(pm, fxy * fxy) + r32(pm, fyy * fyy))));
            double regx = r32(pm, 1.0 / r32(pm, r32(pm, sx * sx) + r32(pm, 1e-6)));
            double regy = r32(pm, 1.0 / r32(pm, r32(pm, sy * sy) + r32(pm, 1e-6)));
            /* reg_x * fxx**2 is reg_x * (fxx * fxx); products left to right */
            J11[q] = (real)r32(pm, r32(pm, regx * r32(pm, fxx * fxx)) + r32(pm, regy * r32(pm, fxy * fxy)));
            J22[q] = (real)r32(pm, r32(pm, regx * r32(pm, fxy * fxy)) + r32(pm, regy * r32(pm, fyy * fyy)));
            J33[q] = (real)r32(pm, r32(pm, regx * r32(pm, fxt * fxt)) + r32(pm, regy * r32(pm, fyt * fyt)));
            J12[q] = (real)r32(pm, r32(pm, r32(pm, regx * fxx) * fxy) + r32(pm, r32(pm, regy * fxy) * fyy));
            J13[q] = (real)r32(pm, r32(pm, r32(pm, regx * fxx) * fxt) + r32(pm, r32(pm, regy * fxy) * fyt));
            J23[q] = (real)r32(pm, r32(pm, r32(pm, regx * fxy) * fxt) + r32(pm, r32(pm, regy * fyy) * fyt));
        }
    t_zero_ring(hp, wp, J11, J22, J33, J12, J13, J23);
    free(F1); free(F2); free(fx); free(ft);
}

/* core/optical_flow.py:163-228 (get_motion_tensor_gray) for one channel. */
void or_motion_tensor_gray(const double *f1, const double *f2, int h, int w, int C, int c,
                           double hy, double hx, int p1, int p2,
                           real *J11, real *J22, real *J33, real *J12, real *J13, real *J23)
{
    const int pm = p1 && p2;
    int hp = h + 2, wp = w + 2;
    double *F1 = t_pad(f1, h, w, C, c), *F2 = t_pad(f2, h, w, C, c);
    for (int j = 1; j <= h; ++j)
        for (int i = 1; i <= w; ++i) {
            size_t q = (size_t)j * wp + i;
            double fx1 = t_grad(p1, F1[q + 1], F1[q - 1], hx), fy1 = t_grad(p1, F1[q + wp], F1[q - wp], hy);
            double fx2 = t_grad(p2, F2[q + 1], F2[q - 1], hx), fy2 = t_grad(p2, F2[q + wp], F2[q - wp], hy);
            double fx = r32(pm, 0.5 * r32(pm, fx1 + fx2));
            double fy = r32(pm, 0.5 * r32(pm, fy1 + fy2));
            double ft = r32(pm, F2[q] - F1[q]);
            J11[q] = (real)r32(pm, fx * fx);
            J22[q] = (real)r32(pm, fy * fy);
            J33[q] = (real)r32(pm, ft * ft);
            J12[q] = (real)r32(pm, fx * fy);
            J13[q] = (real)r32(pm, fx * ft);
            J23[q] = (real)r32(pm, fy * ft);
        }
    t_zero_ring(hp, wp, J11, J22, J33, J12, J13, J23);
    free(F1); free(F2);
}

/* smoothed Heaviside of one directional difference,
 * 0.5 * (1.0 + r / np.sqrt(r * r + eps2)), r = (neighbour - centre) / dist */
static double t_census_s(int p, double nb, double ce, double dist, double eps2)
{
    double r = r32(p, r32(p, nb - ce) / r32(p, dist));
    double d = r32(p, sqrt(r32(p, r32(p, r * r) + r32(p, eps2))));
    return r32(p, 0.5 * r32(p, 1.0 + r32(p, r / d)));
}

/* core/optical_flow.py:231-345 (get_motion_tensor_cs) for one channel.
 * The eight neighbour offsets in row-major order; per offset, the census
 * images s1, s2 on the h x w core, edge-padded; their np.gradient
 * derivatives at the core (the padded ring of J is zeroed at the end). */
void or_motion_tensor_cs(const double *f1, const double *f2, int h, int w, int C, int c,
                         double hy, double hx, double eps, int p1, int p2,
                         real *J11, real *J22, real *J33, real *J12, real *J13, real *J23)
{
    const int pm = p1 && p2;
    int hp = h + 2, wp = w + 2;
    double eps2 = eps * eps;
    double *F1 = t_pad(f1, h, w, C, c), *F2 = t_pad(f2, h, w, C, c);
    double *S1 = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *S2 = (double *)malloc(sizeof(double) * (size_t)h * w);
    double *A[6];
    for (int e = 0; e < 6; ++e) A[e] = (double *)calloc((size_t)h * w, sizeof(double));
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (dy == 0 && dx == 0) continue;
            double dist = sqrt((hy * dy) * (hy * dy) + (hx * dx) * (hx * dx));
            for (int j = 0; j < h; ++j)
                for (int i = 0; i < w; ++i) {
                    size_t qc = (size_t)(j + 1) * wp + (i + 1), qn = qc + (ptrdiff_t)dy * wp + dx;
                    S1[(size_t)j * w + i] = t_census_s(p1, F1[qn], F1[qc], dist, eps2);
                    S2[(size_t)j * w + i] = t_census_s(p2, F2[qn], F2[qc], dist, eps2);
                }
            for (int j = 0; j < h; ++j)
                for (int i = 0; i < w; ++i) {
                    int il = i > 0 ? i - 1 : 0, ir = i < w - 1 ? i + 1 : w - 1;
                    int ju = j > 0 ? j - 1 : 0, jd = j < h - 1 ? j + 1 : h - 1;
                    size_t q = (size_t)j * w + i;
                    double sx1 = t_grad(p1, S1[(size_t)j * w + ir], S1[(size_t)j * w + il], hx);
                    double sx2 = t_grad(p2, S2[(size_t)j * w + ir], S2[(size_t)j * w + il], hx);
                    double sy1 = t_grad(p1, S1[(size_t)jd * w + i], S1[(size_t)ju * w + i], hy);
                    double sy2 = t_grad(p2, S2[(size_t)jd * w + i], S2[(size_t)ju * w + i], hy);
                    double sx = r32(pm, 0.5 * r32(pm, sx1 + sx2));
                    double sy = r32(pm, 0.5 * r32(pm, sy1 + sy2));
                    double st = r32(pm, S2[q] - S1[q]);
                    A[0][q] += r32(pm, sx * sx);
                    A[1][q] += r32(pm, sy * sy);
                    A[2][q] += r32(pm, st * st);
                    A[3][q] += r32(pm, sx * sy);
                    A[4][q] += r32(pm, sx * st);
                    A[5][q] += r32(pm, sy * st);
                }
        }
    real *out[6] = {J11, J22, J33, J12, J13, J23};
    const double scale = 2.0 / 8.0;
    for (int e = 0; e < 6; ++e)
        for (int j = 0; j < h; ++j)
            for (int i = 0; i < w; ++i)
                out[e][(size_t)(j + 1) * wp + (i + 1)] = (real)(A[e][(size_t)j * w + i] * scale);
    t_zero_ring(hp, wp, J11, J22, J33, J12, J13, J23);
    for (int e = 0; e < 6; ++e) free(A[e]);
    free(F1); free(F2); free(S1); free(S2);
}

/* core/level_solver.py:45-66 (set_boundary_2d): rows, then columns. */
static void set_boundary_2d(real *f, int m, int n)
{
    for (int i = 0; i < n; ++i) {
        f[i] = f[(size_t)n + i];
        f[(size_t)(m - 1) * n + i] = f[(size_t)(m - 2) * n + i];
    }
    for (int j = 0; j < m; ++j) {
        f[(size_t)j * n] = f[(size_t)j * n + 1];
        f[(size_t)j * n + n - 1] = f[(size_t)j * n + n - 2];
    }
}

static inline real cf_pow(real x, real b);

/* core/level_solver.py:69-155 (nonlinearity_smoothness_2d), eps = 1e-5,
 * one-sided differences on the outer ring; the power with frh_det_pow, as
 * for the data term's psi (pinned against numpy's pow and np.gradient). */
static void psi_smooth_2d(real *psi_s, const real *u, const real *du, const real *v, const real *dv,
                          int m, int n, real a, real hx, real hy)
{
    size_t N = (size_t)m * n;
    real *uf = (real *)malloc(sizeof(real) * N), *vf = (real *)malloc(sizeof(real) * N);
    for (size_t q = 0; q < N; ++q) { uf[q] = u[q] + du[q]; vf[q] = v[q] + dv[q]; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) {
            size_t q = (size_t)j * n + i;
            real ux = 0, vx = 0, uy = 0, vy = 0;
            if (n > 1) {
                if (i == 0) { ux = (uf[q + 1] - uf[q]) / hx; vx = (vf[q + 1] - vf[q]) / hx; }
                else if (i == n - 1) { ux = (uf[q] - uf[q - 1]) / hx; vx = (vf[q] - vf[q - 1]) / hx; }
                else { ux = (uf[q + 1] - uf[q - 1]) / (2 * hx); vx = (vf[q + 1] - vf[q - 1]) / (2 * hx); }
            }
            if (m > 1) {
                if (j == 0) { uy = (uf[q + n] - uf[q]) / hy; vy = (vf[q + n] - vf[q]) / hy; }
                else if (j == m - 1) { uy = (uf[q] - uf[q - n]) / hy; vy = (vf[q] - vf[q - n]) / hy; }
                else { uy = (uf[q + n] - uf[q - n]) / (2 * hy); vy = (vf[q + n] - vf[q - n]) / (2 * hy); }
            }
            real t = ux * ux + uy * uy + vx * vx + vy * vy;
            if (t < 0) t = 0;
            psi_s[q] = a * cf_pow(t + (real)1e-5, a - 1);
        }
    free(uf); free(vf);
}

/* Exported for the tests: psi_smooth_2d over (m, n) padded arrays. */
void or_psi_smooth(const real *u, const real *du, const real *v, const real *dv, int m, int n,
                   double a, double hx, double hy, real *psi_s)
{
    psi_smooth_2d(psi_s, u, du, v, dv, m, n, (real)a, (real)hx, (real)hy);
}

/* One SOR update of pixel (j, i); core/level_solver.py:262-364. */
static inline void sor_pixel(int j, int i, int m, int n, int K,
                             const real *J11, const real *J22, const real *J12,
                             const real *J13, const real *J23, const real *W,
                             const real *u, const real *v, real *du, real *dv,
                             const real *psi, const real *psi_s, int smooth_const,
                             real ax, real ay, real omega)
{
    size_t c = (size_t)j * n + i;
    size_t nb[4] = {c - 1, c + 1, c + n, c - n}; /* left, right, down, up */
    real wdir[4] = {ax, ax, ay, ay};
    real den_u = 0, den_v = 0, num_u = 0, num_v = 0;
    for (int d = 0; d < 4; ++d) {
        real t = smooth_const ? wdir[d] : (real)0.5 * (psi_s[c] + psi_s[nb[d]]) * wdir[d];
        num_u += t * (u[nb[d]] + du[nb[d]] - u[c]);
        num_v += t * (v[nb[d]] + dv[nb[d]] - v[c]);
        den_u += t;
        den_v += t;
    }
    for (int k = 0; k < K; ++k) {
        size_t ck = c * K + k;
        num_u -= W[ck] * psi[ck] * (J13[ck] + J12[ck] * dv[c]);
        den_u += W[ck] * psi[ck] * J11[ck];
        den_v += W[ck] * psi[ck] * J22[ck];
    }
    real du_k = den_u != 0 ? num_u / den_u : 0;
    du[c] = (1 - omega) * du[c] + omega * du_k;
    real num_v2 = num_v;
    for (int k = 0; k < K; ++k) {
        size_t ck = c * K + k;
        num_v2 -= W[ck] * psi[ck] * (J23[ck] + J12[ck] * du[c]);
    }
    real dv_k = den_v != 0 ? num_v2 / den_v : 0;
    dv[c] = (1 - omega) * dv[c] + omega * dv_k;
}

/* Per-cell linear-system coefficients for one lag block of compute_flow.
 * Restates core/level_solver.py:229-255 (psi refresh) and the constant parts
 * of :262-364 (smoothness weights, data terms) for one pixel c.  The
 * reference recomputes these sums inside every sweep; they only change at a
 * psi refresh, so they are formed once per lag block.  Summation order is
 * fixed here and the GPU kernels use the same order:
 *   weights  wl, wr, wd, wu (left, right, down, up)
 *   Du  = (((wl + wr) + wd) + wu) + sum_k (W_k psi_k) J11_k
 *   Dv  = (((wl + wr) + wd) + wu) + sum_k (W_k psi_k) J22_k
 *   A12 = sum_k (W_k psi_k) J12_k
 *   Bu  = cu - sum_k (W_k psi_k) J13_k,   cu = sum_nb w_nb (u_nb - u_c)
 *   Bv  = cv - sum_k (W_k psi_k) J23_k
 *   coef = {A12, Bu, Bv, omega / Du, omega / Dv}  (0 where D == 0, as the
 *          reference's "num / den if den != 0 else 0")
 * psi_k = a_k (E_k + 1e-5)^(a_k - 1) with E_k the data term at (du, dv),
 * evaluated with frh_det_pow. */
static inline real cf_pow(real x, real b)
{
    if (sizeof(real) == sizeof(float)) return (real)frh_det_powf((float)x, (float)b);
    return (real)frh_det_pow((double)x, (double)b);
}

static const real OMEGA = (real)1.95;

/* core/level_solver.py:411-431 (compute_flow_gnc): a data (or smoothness)
 * exponent strictly between 0 and 1 and a stage weight below 1 blend the
 * robust weight with the quadratic one, (1 - beta) + beta * robust.
 * gnc_beta < 0 selects compute_flow (no GNC). */
static int gnc_uses_mix(double a, double gnc_beta)
{
    return gnc_beta >= 0 && a > 0.0 && a < 1.0 && gnc_beta < 1.0;
}

static void assemble_pixel(size_t c, int n, int K, const real *J11, const real *J22, const real *J33,
                           const real *J12, const real *J13, const real *J23, const real *W,
                           const real *u, const real *v, const real *du, const real *dv,
                           const real *psi_s, int smooth_const, real ax, real ay,
                           const double *a_data, int psi_one, double gnc_beta, real *coef)
{
    size_t nb[4] = {c - 1, c + 1, c + n, c - n};
    real w[4];
    if (smooth_const) {
        w[0] = ax; w[1] = ax; w[2] = ay; w[3] = ay;
    } else {
        real wd[4] = {ax, ax, ay, ay};
        for (int d = 0; d < 4; ++d) w[d] = (real)0.5 * (psi_s[c] + psi_s[nb[d]]) * wd[d];
    }
    real cu = w[0] * (u[nb[0]] - u[c]);
    real cv = w[0] * (v[nb[0]] - v[c]);
    for (int d = 1; d < 4; ++d) {
        cu = cu + w[d] * (u[nb[d]] - u[c]);
        cv = cv + w[d] * (v[nb[d]] - v[c]);
    }
    real ws = ((w[0] + w[1]) + w[2]) + w[3];
    real Du = ws, Dv = ws, A12 = 0, Bu = cu, Bv = cv;
    for (int k = 0; k < K; ++k) {
        size_t ck = c * K + k;
        real psi = 1;
        if (!psi_one) {
            real a = (real)a_data[k];
            real x = du[c], y = dv[c];
            real val = J11[ck] * x * x + J22[ck] * y * y + J23[ck] * y + 2 * J12[ck] * x * y
                     + 2 * J13[ck] * x + J23[ck] * y + J33[ck];
            if (val < 0) val = 0;
            psi = a * cf_pow(val + (real)1e-5, a - 1);
            if (gnc_uses_mix(a_data[k], gnc_beta))
                psi = (real)(1.0 - gnc_beta) + (real)gnc_beta * psi;
        }
        real wp = W[ck] * psi;
        Du = Du + wp * J11[ck];
        Dv = Dv + wp * J22[ck];
        A12 = A12 + wp * J12[ck];
        Bu = Bu - wp * J13[ck];
        Bv = Bv - wp * J23[ck];
    }
    coef[0] = A12; coef[1] = Bu; coef[2] = Bv;
    coef[3] = Du != 0 ? OMEGA / Du : 0;
    coef[4] = Dv != 0 ? OMEGA / Dv : 0;
    coef[5] = w[0]; coef[6] = w[1]; coef[7] = w[2]; coef[8] = w[3];
}

/* One SOR update with the lag-block coefficients; core/level_solver.py:
 * 342-364.  A neighbour outside the interior holds the pixel's own value
 * from the start of the sweep (set_boundary_2d, :257-258, copies the
 * adjacent interior value into the ring before every sweep, and u is
 * edge-padded), so its weight moves to the pixel's own term:
 *   du' = (1 - omega) du + (omega / Du) (Bu + sum_in w_nb du_nb
 *          + b du - A12 dv),   b = sum of out-of-level weights
 *       = fma(wu, n, fma(wu, b, 1 - omega) du)
 *   n   = fma(-A12, dv, fma(wd_sum...)), in the order below (shared with the
 *         GPU kernel). */
static inline void sor_pixel_coef(int j, int i, int m, int n, real *du, real *dv,
                                  const real *coef, int smooth_const)
{
    size_t c = (size_t)j * n + i;
    int in_l = i > 1, in_r = i < n - 2, in_d = j < m - 2, in_u = j > 1;
    real dul = in_l ? du[c - 1] : 0, dur = in_r ? du[c + 1] : 0;
    real dud = in_d ? du[c + n] : 0, duu = in_u ? du[c - n] : 0;
    real dvl = in_l ? dv[c - 1] : 0, dvr = in_r ? dv[c + 1] : 0;
    real dvd = in_d ? dv[c + n] : 0, dvu = in_u ? dv[c - n] : 0;
    real A12 = coef[0], Bu = coef[1], Bv = coef[2], wu = coef[3], wv = coef[4];
    real wl = coef[5], wr = coef[6], wd = coef[7], wup = coef[8];
    real b;
    if (smooth_const) /* ax * (#left/right outside) + ay * (#up/down outside) */
        b = fma(wl, (real)(!in_l + !in_r), wd * (real)(!in_u + !in_d));
    else
        b = (((in_l ? 0 : wl) + (in_r ? 0 : wr)) + (in_u ? 0 : wup)) + (in_d ? 0 : wd);
    real ou = du[c], ov = dv[c];
    const real om1 = 1 - OMEGA;
    real nu, nv;
    if (smooth_const) {
        nu = fma(wl, dul + dur, Bu);
        nu = fma(wd, dud + duu, nu);
        nv = fma(wl, dvl + dvr, Bv);
        nv = fma(wd, dvd + dvu, nv);
    } else {
        nu = fma(wup, duu, fma(wd, dud, fma(wr, dur, fma(wl, dul, Bu))));
        nv = fma(wup, dvu, fma(wd, dvd, fma(wr, dvr, fma(wl, dvl, Bv))));
    }
    nu = fma(-A12, ov, nu);
    real dun = fma(wu, nu, fma(wu, b, om1) * ou);
    nv = fma(-A12, dun, nv);
    du[c] = dun;
    dv[c] = fma(wv, nv, fma(wv, b, om1) * ov);
}

/* core/level_solver.py:158-369 (compute_flow).
 * J*, W: (m, n, K); u, v: (m, n); du, dv: (m, n) outputs.
 * order: 0 = column-major lexicographic sweep with the lag-block coefficient
 *            form above (the parity oracle; the GPU matches it bit for bit);
 *        1 = the same with red-black ordering (study only);
 *        2 = the reference's per-sweep expression order with libm pow
 *            (study only: shows the rounding sensitivity of the method). */
void or_compute_flow(const real *J11, const real *J22, const real *J33, const real *J12,
                     const real *J13, const real *J23, const real *W,
                     const real *u, const real *v, int m, int n, int K,
                     double alpha_x, double alpha_y, int iterations, int update_lag,
                     const double *a_data, double a_smooth, double hx, double hy,
                     double gnc_beta, real *du, real *dv, int order)
{
    size_t N = (size_t)m * n;
    real *psi = (real *)malloc(sizeof(real) * N * K);
    real *psi_s = (real *)malloc(sizeof(real) * N);
    real *coef = (real *)malloc(sizeof(real) * N * 9);
    for (size_t q = 0; q < N * K; ++q) psi[q] = 1;
    for (size_t q = 0; q < N; ++q) { psi_s[q] = 1; du[q] = 0; dv[q] = 0; }
    const real omega = OMEGA;
    real ax = (real)(alpha_x / (hx * hx)), ay = (real)(alpha_y / (hy * hy));
    int smooth_const = (a_smooth == 1.0);
    for (int it = 0; it < iterations; ++it) {
        int refresh = ((it + 1) % update_lag == 0);
        if (refresh) {
            if (order == 2)
                for (int k = 0; k < K; ++k)
                    for (size_t q = 0; q < N; ++q) {
                        size_t qk = q * K + k;
                        real val = J11[qk] * du[q] * du[q] + J22[qk] * dv[q] * dv[q] + J23[qk] * dv[q]
                                 + 2 * J12[qk] * du[q] * dv[q] + 2 * J13[qk] * du[q] + J23[qk] * dv[q]
                                 + J33[qk];
                        if (val < 0) val = 0;
                        psi[qk] = (real)a_data[k] * pow(val + (real)1e-5, (real)a_data[k] - 1);
                        if (gnc_uses_mix(a_data[k], gnc_beta))
                            psi[qk] = (real)(1.0 - gnc_beta) + (real)gnc_beta * psi[qk];
                    }
            if (!smooth_const) {
                psi_smooth_2d(psi_s, u, du, v, dv, m, n, (real)a_smooth, (real)hx, (real)hy);
                if (gnc_uses_mix(a_smooth, gnc_beta))
                    for (size_t q = 0; q < N; ++q)
                        psi_s[q] = (real)(1.0 - gnc_beta) + (real)gnc_beta * psi_s[q];
            } else
                for (size_t q = 0; q < N; ++q) psi_s[q] = 1;
        }
        if (order != 2 && (refresh || it == 0))
            for (int i = 1; i < n - 1; ++i)
                for (int j = 1; j < m - 1; ++j) {
                    size_t c = (size_t)j * n + i;
                    assemble_pixel(c, n, K, J11, J22, J33, J12, J13, J23, W, u, v, du, dv, psi_s,
                                   smooth_const, ax, ay, a_data, !refresh, gnc_beta, coef + c * 9);
                }
        set_boundary_2d(du, m, n);
        set_boundary_2d(dv, m, n);
        if (order == 2) {
            for (int i = 1; i < n - 1; ++i)
                for (int j = 1; j < m - 1; ++j)
                    sor_pixel(j, i, m, n, K, J11, J22, J12, J13, J23, W, u, v, du, dv,
                              psi, psi_s, smooth_const, ax, ay, omega);
        } else if (order == 0) {
            for (int i = 1; i < n - 1; ++i)
                for (int j = 1; j < m - 1; ++j)
                    sor_pixel_coef(j, i, m, n, du, dv, coef + ((size_t)j * n + i) * 9, smooth_const);
        } else {
            for (int colour = 0; colour < 2; ++colour)
                for (int i = 1; i < n - 1; ++i)
                    for (int j = 1; j < m - 1; ++j)
                        if (((i + j) & 1) == colour)
                            sor_pixel_coef(j, i, m, n, du, dv, coef + ((size_t)j * n + i) * 9, smooth_const);
        }
    }
    free(psi); free(psi_s); free(coef);
}

/* cv2.resize(w, (n, m), interpolation=INTER_CUBIC) for float images
 * (core/optical_flow.py:615-616, only when min_level > 0).  OpenCV's public
 * behaviour: source coordinate (i+0.5)/scale-0.5, Keys cubic A=-0.75 with
 * float coefficients, replicated borders, horizontal pass then vertical.
 * Parity with OpenCV itself is unpinned. */
void or_cubic_resize_cv(const real *src, int h, int w, int C, real *dst, int oh, int ow)
{
    double scale_x = 1.0 / ((double)ow / w), scale_y = 1.0 / ((double)oh / h);
    real *tmp = (real *)malloc(sizeof(real) * (size_t)h * ow * C);
    for (int x = 0; x < ow; ++x) {
        float f = (float)((x + 0.5) * scale_x - 0.5);
        int ix = (int)floorf(f);
        float t = f - (float)ix, cf[4];
        cv_cubic_coeffs(t, cf);
        for (int y = 0; y < h; ++y)
            for (int c = 0; c < C; ++c) {
                real s = 0;
                for (int k = 0; k < 4; ++k) {
                    int xx = ix - 1 + k;
                    xx = xx < 0 ? 0 : (xx >= w ? w - 1 : xx);
                    s += src[((size_t)y * w + xx) * C + c] * (real)cf[k];
                }
                tmp[((size_t)y * ow + x) * C + c] = s;
            }
    }
    for (int y = 0; y < oh; ++y) {
        float f = (float)((y + 0.5) * scale_y - 0.5);
        int iy = (int)floorf(f);
        float t = f - (float)iy, cf[4];
        cv_cubic_coeffs(t, cf);
        for (int x = 0; x < ow; ++x)
            for (int c = 0; c < C; ++c) {
                real s = 0;
                for (int k = 0; k < 4; ++k) {
                    int yy = iy - 1 + k;
                    yy = yy < 0 ? 0 : (yy >= h ? h - 1 : yy);
                    s += tmp[((size_t)yy * ow + x) * C + c] * (real)cf[k];
                }
                dst[((size_t)y * ow + x) * C + c] = s;
            }
    }
    free(tmp);
}
