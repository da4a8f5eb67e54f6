// HIP kernels for the variational optical-flow path on gfx950.
//
// Every kernel reproduces the arithmetic of one reference stage in a fixed
// operation order (compiled with -ffp-contract=off, IEEE division and
// square root) so that results are bit-identical to the CPU oracle given the
// same inputs.  Reference citations are pyflowreg source paths.
#include <type_traits>
#include <utility>

#include "frh_internal.h"
#include "frh_device.h"
#include "frh_median.h"
#include "../../include/flowreg_detmath.h"

namespace frh {

__device__ inline int reflect_half(int j, int n)  // ...b a | a b...  (scipy "reflect")
{
    if (n <= 1) return 0;
    while (j < 0 || j >= n) j = j < 0 ? -j - 1 : 2 * n - 1 - j;
    return j;
}

// Work-groups are dispatched to the 8 XCDs of an MI355X round-robin (group b
// runs on XCD b % 8).  Renumbering them so that each XCD receives one
// contiguous range of the nb logical groups keeps neighbouring tiles in the
// same L2.  A bijection of [0, nb).
__device__ inline unsigned xcd_block(unsigned b, unsigned nb)
{
    const unsigned X = 8, per = nb / X, rem = nb % X;
    const unsigned x = b % X, k = b / X;
    return (x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per) + k;
}

// The logical block (x, y, z) of a gx x gy x gz grid launched as
// gx * gy * gz one-dimensional blocks (x fastest, then y, then z), numbered
// with xcd_block: each XCD runs a contiguous range, i.e. whole frames, so
// rows that the taps of neighbouring output rows share stay in one L2.
struct Blk3 {
    int x, y, z;
};

__device__ inline Blk3 xcd_block3(int gx, int gy)
{
    const unsigned gxy = (unsigned)gx * (unsigned)gy;
    const unsigned L = xcd_block(blockIdx.x, gridDim.x);
    return Blk3{(int)(L % (unsigned)gx), (int)((L / (unsigned)gx) % (unsigned)gy), (int)(L / gxy)};
}

__device__ inline int mirror_edge(int j, int n)  // ...c b | a b c...  (scipy "mirror")
{
    if (n >= 3 && j >= -(n - 1) && j <= 2 * n - 2) {  // one reflection suffices
        j = j < 0 ? -j : j;
        return j >= n ? 2 * n - 2 - j : j;
    }
    if (n == 1) return 0;
    int period = 2 * n - 2;
    j %= period;
    if (j < 0) j += period;
    return j < n ? j : period - j;
}

__device__ inline int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// -------------------------------------------------------------- prefilter
// util/image_processing.py:12-98 + :101-165.  scipy.ndimage.correlate1d with
// a symmetric kernel sums  x[0] w[0] + sum_{j=r..1} (x[-j] + x[j]) w[j].
struct GaussW {
    double w[65];
};

// sem (FRH_PRE_* of include/flowreg_hip.h): bit 0, the normalisation
// (x - lo) / den in float32 arithmetic (numpy with a float32 array and
// float32 bounds: each operation rounded to float32, which double
// arithmetic followed by a float32 rounding reproduces exactly); bit 1,
// scipy's gaussian_filter on a float32 array, which stores each 1-D pass
// into the float32 output (the double line sums rounded to float32).
constexpr int kPreNorm32 = 1, kPreFilt32 = 2;

// Column pass: a work-group of 64 lanes takes kPreRows output rows of 64
// columns; each lane normalises the input rows its outputs need once, into
// its own column of LDS (no lane reads another's, so no barrier), and forms
// each output from them with the operations and order of the row sum below.
constexpr int kPreRows = 16;

template <typename T>
__global__ __launch_bounds__(64) void k_prefilter_v(const T *__restrict__ raw, int H, int W, int C, GaussW lo_den,
                                                   GaussW wy, int ry, int sem, double *__restrict__ tmp)
{
    extern __shared__ double win[];  // (kPreRows + 2 ry) rows x 64 lanes
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y0 = blockIdx.y * kPreRows, n = H - y0 < kPreRows ? H - y0 : kPreRows;
    const int tc = blockIdx.z;
    if (x >= W) return;
    const int t = tc / C, c = tc % C;
    const double lo = lo_den.w[2 * c], den = lo_den.w[2 * c + 1];
    const T *base = raw + (size_t)t * H * W * C + c;
    const bool n32 = (sem & kPreNorm32) != 0;
    auto nrm = [&](int yy) {
        const double d = (double)base[((size_t)yy * W + x) * C] - lo;
        return n32 ? (double)(float)((double)(float)d / den) : d / den;
    };
    // row k of the window is input row y0 - ry + k (reflected into [0, H))
    double *col = win + threadIdx.x;
    for (int k = 0; k < n + 2 * ry; ++k) col[k * 64] = nrm(reflect_half(y0 - ry + k, H));
    for (int i = 0; i < n; ++i) {
        const double *cw = col + (i + ry) * 64;  // output row y0 + i
        double acc = cw[0] * wy.w[ry];
        for (int j = ry; j >= 1; --j) acc += (cw[-j * 64] + cw[j * 64]) * wy.w[ry - j];
        tmp[((size_t)tc * H + y0 + i) * W + x] = (sem & kPreFilt32) ? (double)(float)acc : acc;
    }
}

// Temporal pass of apply_gaussian_filter's 4-D branch (scipy's
// gaussian_filter filters axis 0, the frames of the batch, first): the
// normalised frames (x - lo) / den correlated with the temporal kernel
// along T with mode "reflect", in the operation order of the row pass;
// float64 (T, H, W, C) out.  One thread per element of a frame row.
template <typename T>
__global__ void k_prefilter_t(const T *__restrict__ raw, int Tn, int H, int W, int C, GaussW lo_den, GaussW wt,
                              int rt, double *__restrict__ out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    const int WC = W * C;
    if (e >= WC) return;
    const int c = e % C;
    const double lo = lo_den.w[2 * c], den = lo_den.w[2 * c + 1];
    const size_t frame = (size_t)H * WC, off = (size_t)y * WC + e;
    auto nrm = [&](int tt) { return ((double)raw[(size_t)tt * frame + off] - lo) / den; };
    double acc = nrm(t) * wt.w[rt];
    for (int j = rt; j >= 1; --j)
        acc += (nrm(reflect_half(t - j, Tn)) + nrm(reflect_half(t + j, Tn))) * wt.w[rt - j];
    out[(size_t)t * frame + off] = acc;
}

hipError_t launch_prefilter_t(const void *raw, int raw_f64, int Tn, int H, int W, int C, const double *lo,
                              const double *den, const double *wt, int rt, double *out, hipStream_t st)
{
    GaussW ld{}, g{};
    for (int c = 0; c < C; ++c) {
        ld.w[2 * c] = lo[c];
        ld.w[2 * c + 1] = den[c];
    }
    for (int k = 0; k < 2 * rt + 1; ++k) g.w[k] = wt[k];
    const dim3 grid((W * C + 63) / 64, H, Tn);
    if (raw_f64)
        hipLaunchKernelGGL(k_prefilter_t<double>, grid, dim3(64), 0, st, (const double *)raw, Tn, H, W, C, ld, g,
                           rt, out);
    else
        hipLaunchKernelGGL(k_prefilter_t<float>, grid, dim3(64), 0, st, (const float *)raw, Tn, H, W, C, ld, g,
                           rt, out);
    return hipGetLastError();
}

// Row pass: the work-group's 64 outputs read their 64 + 2 rx inputs
// (reflected at the row's ends) from LDS, loaded once.
template <typename O>
__global__ __launch_bounds__(64) void k_prefilter_h(const double *__restrict__ tmp, int H, int W, GaussW wx,
                                                   int rx, int sem, O *__restrict__ out)
{
    extern __shared__ double seg[];  // inputs x0 - rx .. x0 + 63 + rx
    const int x0 = blockIdx.x * blockDim.x, x = x0 + threadIdx.x;
    const int y = blockIdx.y, tc = blockIdx.z;
    const double *row = tmp + ((size_t)tc * H + y) * W;
    const int nx = W - x0 < 64 ? W - x0 : 64;  // outputs of this work-group
    for (int k = threadIdx.x; k < nx + 2 * rx; k += blockDim.x) seg[k] = row[reflect_half(x0 - rx + k, W)];
    __syncthreads();
    if (x >= W) return;
    const double *cw = seg + threadIdx.x + rx;
    double acc = cw[0] * wx.w[rx];
    for (int j = rx; j >= 1; --j) acc += (cw[-j] + cw[j]) * wx.w[rx - j];
    out[((size_t)tc * H + y) * W + x] = (sem & kPreFilt32) ? (O)(float)acc : (O)acc;
}

hipError_t launch_prefilter_v(const float *raw, int T, int H, int W, int C, const double *lo,
                              const double *den, const double *wy, int ry, int sem, double *tmp,
                              hipStream_t st)
{
    GaussW ld{}, g{};
    for (int c = 0; c < C; ++c) {
        ld.w[2 * c] = lo[c];
        ld.w[2 * c + 1] = den[c];
    }
    for (int k = 0; k < 2 * ry + 1; ++k) g.w[k] = wy[k];
    dim3 grid((W + 63) / 64, (H + kPreRows - 1) / kPreRows, T * C);
    hipLaunchKernelGGL(k_prefilter_v<float>, grid, dim3(64), sizeof(double) * 64 * (kPreRows + 2 * ry), st, raw, H,
                       W, C, ld, g, ry, sem, tmp);
    return hipGetLastError();
}

hipError_t launch_prefilter_v64(const double *raw, int T, int H, int W, int C, const double *lo,
                                const double *den, const double *wy, int ry, int sem, double *tmp,
                                hipStream_t st)
{
    GaussW ld{}, g{};
    for (int c = 0; c < C; ++c) {
        ld.w[2 * c] = lo[c];
        ld.w[2 * c + 1] = den[c];
    }
    for (int k = 0; k < 2 * ry + 1; ++k) g.w[k] = wy[k];
    dim3 grid((W + 63) / 64, (H + kPreRows - 1) / kPreRows, T * C);
    hipLaunchKernelGGL(k_prefilter_v<double>, grid, dim3(64), sizeof(double) * 64 * (kPreRows + 2 * ry), st, raw,
                       H, W, C, ld, g, ry, sem, tmp);
    return hipGetLastError();
}

hipError_t launch_prefilter_h(const double *tmp, int T, int H, int W, int C, const double *wx,
                              int rx, int sem, float *out, hipStream_t st)
{
    GaussW g{};
    for (int k = 0; k < 2 * rx + 1; ++k) g.w[k] = wx[k];
    dim3 grid((W + 63) / 64, H, T * C);
    hipLaunchKernelGGL(k_prefilter_h<float>, grid, dim3(64), sizeof(double) * (64 + 2 * rx), st, tmp, H, W, g, rx,
                       sem, out);
    return hipGetLastError();
}

hipError_t launch_prefilter_h64(const double *tmp, int T, int H, int W, int C, const double *wx,
                                int rx, int sem, double *out, hipStream_t st)
{
    GaussW g{};
    for (int k = 0; k < 2 * rx + 1; ++k) g.w[k] = wx[k];
    dim3 grid((W + 63) / 64, H, T * C);
    hipLaunchKernelGGL(k_prefilter_h<double>, grid, dim3(64), sizeof(double) * (64 + 2 * rx), st, tmp, H, W, g, rx,
                       sem, out);
    return hipGetLastError();
}

// ----------------------------------------------------------------- resize
// util/resize_util.py:116-141: float32 products, float64 running sums,
// float32 stores; horizontal pass then vertical pass.

// Sum over the P taps of a table row: the float32 product of the tap's
// value (load(index)) and weight, added in tap order to a float64 running
// sum.  The table entries and values of eight taps are loaded before any of
// their products is added, so eight memory round trips overlap instead of
// following one another; the arithmetic and its order are unchanged.
template <typename Load>
__device__ __forceinline__ double tap_sum(int P, const int32_t *ii, const float *ww, Load load)
{
    double s = 0.0;
    int p = 0;
    for (; p + 8 <= P; p += 8) {
        float v[8], wt[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            v[k] = load(ii[p + k]);
            wt[k] = ww[p + k];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float prod = v[k] * wt[k];
            s += (double)prod;
        }
    }
    for (; p < P; ++p) {
        const float prod = load(ii[p]) * ww[p];
        s += (double)prod;
    }
    return s;
}

__global__ void k_resize_h(const float *__restrict__ src, int h, int w, float *__restrict__ tmp,
                           int ow, const int32_t *__restrict__ ix, const float *__restrict__ wx,
                           int P)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    int n = blockIdx.z;
    if (x >= ow) return;
    const float *row = src + ((size_t)n * h + y) * w;
    const double s = tap_sum(P, ix + (size_t)x * P, wx + (size_t)x * P, [&](int j) { return row[j]; });
    tmp[((size_t)n * h + y) * ow + x] = (float)s;
}

__global__ void k_resize_v(const float *__restrict__ tmp, int h, float *__restrict__ dst, int oh,
                           int ow, const int32_t *__restrict__ iy, const float *__restrict__ wy,
                           int P)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y;
    int n = blockIdx.z;
    if (x >= ow) return;
    const float *col = tmp + (size_t)n * h * ow + x;
    const double s = tap_sum(P, iy + (size_t)y * P, wy + (size_t)y * P,
                             [&](int i) { return col[(size_t)i * ow]; });
    dst[((size_t)n * oh + y) * ow + x] = (float)s;
}

hipError_t launch_resize(const float *src, int N, int h, int w, float *dst, int oh, int ow,
                         const int32_t *ix, const float *wx, const int32_t *iy, const float *wy,
                         int P, float *tmp, hipStream_t st)
{
    hipLaunchKernelGGL(k_resize_h, dim3((ow + 63) / 64, h, N), dim3(64), 0, st, src, h, w, tmp, ow,
                       ix, wx, P);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_resize_v, dim3((ow + 63) / 64, oh, N), dim3(64), 0, st, tmp, h, dst, oh,
                       ow, iy, wy, P);
    return hipGetLastError();
}

// Both passes in one kernel for short tap lists (the flow upsampling
// between levels, P = 4): each output sample forms the horizontal sums of
// the P rows its vertical taps name -- each the same float32 products and
// float64 running sum as k_resize_h, rounded to float32 -- and combines them
// as k_resize_v does, so the result is the two-pass result bit for bit,
// without the intermediate plane.
// Output sample (x, y) of one plane of width w.
template <int P>
__device__ __forceinline__ float resize_fused_px(const float *__restrict__ plane, int w,
                                                 const int32_t *__restrict__ ix, const float *__restrict__ wx,
                                                 const int32_t *__restrict__ iy, const float *__restrict__ wy,
                                                 int x, int y)
{
    int cx[P];
    float kx[P];
#pragma unroll
    for (int q = 0; q < P; ++q) {
        cx[q] = ix[(size_t)x * P + q];
        kx[q] = wx[(size_t)x * P + q];
    }
    double s = 0.0;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const float *row = plane + (size_t)iy[(size_t)y * P + p] * w;
        double hs = 0.0;
#pragma unroll
        for (int q = 0; q < P; ++q) {
            const float prod = row[cx[q]] * kx[q];
            hs += (double)prod;
        }
        const float prod = (float)hs * wy[(size_t)y * P + p];
        s += (double)prod;
    }
    return (float)s;
}

template <int P>
__global__ void k_resize_fused(const float *__restrict__ src, int h, int w, float *__restrict__ dst, int oh,
                               int ow, const int32_t *__restrict__ ix, const float *__restrict__ wx,
                               const int32_t *__restrict__ iy, const float *__restrict__ wy)
{
    const Blk3 b = xcd_block3((ow + blockDim.x - 1) / blockDim.x, oh);
    const int x = b.x * blockDim.x + threadIdx.x, y = b.y, n = b.z;
    if (x >= ow) return;
    dst[((size_t)n * oh + y) * ow + x] = resize_fused_px<P>(src + (size_t)n * h * w, w, ix, wx, iy, wy, x, y);
}

hipError_t launch_resize_fused4(const float *src, int N, int h, int w, float *dst, int oh, int ow,
                                const int32_t *ix, const float *wx, const int32_t *iy, const float *wy,
                                hipStream_t st)
{
    const size_t nb = (size_t)((ow + 63) / 64) * oh * N;
    if (nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_resize_fused<4>, dim3((unsigned)nb), dim3(64), 0, st, src, h, w, dst, oh, ow, ix, wx, iy,
                       wy);
    return hipGetLastError();
}

// Multi-level pyramid from full resolution (core/optical_flow.py:519-520:
// every level is resized from the full-resolution frame).  Horizontal pass:
// one work-group per (plane, block of R rows) stages the rows in LDS and
// emits their output samples for every level at once (outputs of all levels
// concatenated, Wsum = sum of level widths; per-output tap lists in hoff /
// hix / hwx, each tap applied to the R rows so the tables are read once per
// R rows).  Vertical pass: one launch per level over the staged rows.
// Same float32 products and float64 sums as k_resize_h / k_resize_v.
template <int R>
__global__ void k_pyr_h(const float *__restrict__ src, int H, int W, int Wsum,
                        const int32_t *__restrict__ hoff, const int32_t *__restrict__ hix,
                        const float *__restrict__ hwx, float *__restrict__ tmp)
{
    extern __shared__ float rows[];  // R staged full-res rows
    const int y0 = blockIdx.x * R, n = blockIdx.y;
    const int nr = H - y0 < R ? H - y0 : R;
    const float *s = src + ((size_t)n * H + y0) * W;
    for (int i = threadIdx.x; i < nr * W; i += blockDim.x) rows[i] = s[i];
    __syncthreads();
    float *o = tmp + ((size_t)n * H + y0) * Wsum;
    for (int x = threadIdx.x; x < Wsum; x += blockDim.x) {
        const int b = hoff[x], e = hoff[x + 1];
        double acc[R];
#pragma unroll
        for (int i = 0; i < R; ++i) acc[i] = 0.0;
        // the table entries of eight taps are loaded before their products
        // are added (each row's sum keeps the tap order)
        int k = b;
        for (; k + 8 <= e; k += 8) {
            int jj[8];
            float wt[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                jj[q] = hix[k + q];
                wt[q] = hwx[k + q];
            }
#pragma unroll
            for (int q = 0; q < 8; ++q)
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    float prod = rows[i * W + jj[q]] * wt[q];
                    acc[i] += (double)prod;
                }
        }
        for (; k < e; ++k) {
            const int j = hix[k];
            const float wt = hwx[k];
#pragma unroll
            for (int i = 0; i < R; ++i) {
                float prod = rows[i * W + j] * wt;
                acc[i] += (double)prod;
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (i < nr) o[(size_t)i * Wsum + x] = (float)acc[i];
    }
}

__global__ void k_pyr_v(const float *__restrict__ tmp, int H, int Wsum, int off,
                        float *__restrict__ dst, int oh, int ow, const int32_t *__restrict__ iy,
                        const float *__restrict__ wy, int P)
{
    const Blk3 b = xcd_block3((ow + blockDim.x - 1) / blockDim.x, oh);
    const int x = b.x * blockDim.x + threadIdx.x, y = b.y, n = b.z;
    if (x >= ow) return;
    const float *col = tmp + (size_t)n * H * Wsum + off + x;
    const double acc = tap_sum(P, iy + (size_t)y * P, wy + (size_t)y * P,
                               [&](int i) { return col[(size_t)i * Wsum]; });
    dst[((size_t)n * oh + y) * ow + x] = (float)acc;
}

hipError_t launch_pyr_h(const float *src, int N, int H, int W, int Wsum, const int32_t *hoff,
                        const int32_t *hix, const float *hwx, float *tmp, hipStream_t st)
{
    if (W <= 1024) {
        constexpr int R = 16;
        hipLaunchKernelGGL(k_pyr_h<R>, dim3((H + R - 1) / R, N), dim3(256), sizeof(float) * R * W, st,
                           src, H, W, Wsum, hoff, hix, hwx, tmp);
    } else {
        constexpr int R = 4;
        hipLaunchKernelGGL(k_pyr_h<R>, dim3((H + R - 1) / R, N), dim3(256), sizeof(float) * R * W, st,
                           src, H, W, Wsum, hoff, hix, hwx, tmp);
    }
    return hipGetLastError();
}

hipError_t launch_pyr_v(const float *tmp, int N, int H, int Wsum, int off, float *dst, int oh,
                        int ow, const int32_t *iy, const float *wy, int P, hipStream_t st)
{
    int bx = ow >= 64 ? 64 : (ow >= 32 ? 32 : 16);
    const size_t nb = (size_t)((ow + bx - 1) / bx) * oh * N;
    if (nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_pyr_v, dim3((unsigned)nb), dim3(bx), 0, st, tmp, H, Wsum, off, dst, oh, ow, iy, wy, P);
    return hipGetLastError();
}

// ------------------------------------------------------------------- warp
// core/warping.py:160-234 with OpenCV remap INTER_CUBIC/BORDER_REPLICATE
// semantics: float maps rounded to 1/32 px, float 1-D Keys coefficients
// (A = -0.75), 4x4 weights as their products, float sums (interior: row by
// row; near the border: the 16 taps in sequence on clamped indices).
__device__ inline void cv_cubic_coeffs(float x, float *c)
{
    const float A = -0.75f;
    c[0] = ((A * (x + 1) - 5 * A) * (x + 1) + 8 * A) * (x + 1) - 4 * A;
    c[1] = ((A + 2) * x - (A + 3)) * x * x + 1;
    c[2] = ((A + 2) * (1 - x) - (A + 3)) * (1 - x) * (1 - x) + 1;
    c[3] = 1.f - c[0] - c[1] - c[2];
}

// OpenCV's 16U / 16S remap paths use the float weights and float sums of
// its 32F path on the integer samples (exact in float) and store
// saturate_cast<T>(sum): round half to even, then clamp to T's range.  With
// sat set, an in-bounds result is rounded and clamped to [lo, hi] that way;
// the out-of-bounds fill is stored as it is (the reference writes it into its
// float32 output array).
__device__ __forceinline__ float warp_store_value(float s, int sat, float lo, float hi)
{
    return sat ? fminf(fmaxf(rintf(s), lo), hi) : s;
}

// U: float (level flows, divided by h_col / h_row in float64 here) or
// double (displacements already divided by the caller, h = 1).
// cv2.remap INTER_CUBIC + BORDER_REPLICATE of the image type I: the 1/32-px
// float weight table of both depths, and sums in Acc -- float for float32
// images (OpenCV's 32F path), double for float64 images (its 64F path, which
// the reference takes for its float64 level images); the result is stored as
// float32, as the reference's float32 output array does.
// One output pixel (x, y) of one frame: F2, F1 and O point at the frame's
// C planes of H x W, (ux, vy) is the pixel's displacement.
template <typename U, typename I, typename Acc>
__device__ __forceinline__ void warp_cubic_px(const I *__restrict__ F2, const I *__restrict__ F1,
                                              float *__restrict__ O, int C, int H, int W, int x, int y,
                                              U ux, U vy, double h_col, double h_row, int sat, float lo,
                                              float hi)
{
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    float mx = (float)((double)x + (double)ux / h_col);
    float my = (float)((double)y + (double)vy / h_row);
    bool oob = (mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H);
    if (oob) {
        for (int c = 0; c < C; ++c) O[c * HW + p] = (float)F1[c * HW + p];
        return;
    }
    float cx = mx > (float)(W - 1) ? (float)(W - 1) : mx;
    float cy = my > (float)(H - 1) ? (float)(H - 1) : my;
    int X = (int)rintf(cx * 32.f), Y = (int)rintf(cy * 32.f);
    int sx = (X >> 5) - 1, sy = (Y >> 5) - 1;
    float wxv[4], wyv[4], w[16];
    cv_cubic_coeffs((float)(X & 31) * (1.f / 32), wxv);
    cv_cubic_coeffs((float)(Y & 31) * (1.f / 32), wyv);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) w[i * 4 + j] = wyv[i] * wxv[j];
    int width1 = W - 3 > 0 ? W - 3 : 0, height1 = H - 3 > 0 ? H - 3 : 0;
    if ((unsigned)sx < (unsigned)width1 && (unsigned)sy < (unsigned)height1) {
        // OpenCV's order: the four products of a row summed left to right,
        // the rows added top to bottom
        auto sum16 = [&](const I *a) {
            Acc sum = (Acc)a[0] * w[0] + (Acc)a[1] * w[1] + (Acc)a[2] * w[2] + (Acc)a[3] * w[3];
            sum += (Acc)a[4] * w[4] + (Acc)a[5] * w[5] + (Acc)a[6] * w[6] + (Acc)a[7] * w[7];
            sum += (Acc)a[8] * w[8] + (Acc)a[9] * w[9] + (Acc)a[10] * w[10] + (Acc)a[11] * w[11];
            sum += (Acc)a[12] * w[12] + (Acc)a[13] * w[13] + (Acc)a[14] * w[14] + (Acc)a[15] * w[15];
            return sum;
        };
        auto load16 = [&](int c, I *a) {
            const I *S = F2 + c * HW + (size_t)sy * W + sx;
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) a[i * 4 + j] = S[(size_t)i * W + j];
        };
        int c = 0;
        // two channels' 32 samples are loaded before either sum is formed
        for (; c + 2 <= C; c += 2) {
            I a0[16], a1[16];
            load16(c, a0);
            load16(c + 1, a1);
            O[c * HW + p] = warp_store_value((float)sum16(a0), sat, lo, hi);
            O[(c + 1) * HW + p] = warp_store_value((float)sum16(a1), sat, lo, hi);
        }
        for (; c < C; ++c) {
            I a0[16];
            load16(c, a0);
            O[c * HW + p] = warp_store_value((float)sum16(a0), sat, lo, hi);
        }
    } else {
        int xs[4], ys[4];
        for (int i = 0; i < 4; ++i) {
            xs[i] = clampi(sx + i, 0, W - 1);
            ys[i] = clampi(sy + i, 0, H - 1);
        }
        for (int c = 0; c < C; ++c) {
            const I *S = F2 + c * HW;
            Acc sum = 0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) sum += (Acc)S[(size_t)ys[i] * W + xs[j]] * w[i * 4 + j];
            O[c * HW + p] = warp_store_value((float)sum, sat, lo, hi);
        }
    }
}

template <typename U, typename I, typename Acc>
__global__ void k_warp(const I *__restrict__ f2, const I *__restrict__ f1,
                       int f1_per_frame, int C, int H, int W, const U *__restrict__ u,
                       const U *__restrict__ v, size_t uv_stride, double h_col, double h_row,
                       float *__restrict__ out, int sat = 0, float lo = 0.f, float hi = 0.f)
{
    const Blk3 b = xcd_block3((W + blockDim.x - 1) / blockDim.x, H);
    const int x = b.x * blockDim.x + threadIdx.x, y = b.y, t = b.z;
    if (x >= W) return;
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    warp_cubic_px<U, I, Acc>(f2 + (size_t)t * C * HW, f1 + (f1_per_frame ? (size_t)t * C * HW : 0),
                             out + (size_t)t * C * HW, C, H, W, x, y, u[t * uv_stride + p],
                             v[t * uv_stride + p], h_col, h_row, sat, lo, hi);
}

// k_warp's one-dimensional grid of (W / 64) x H x T blocks (xcd_block3)
static dim3 warp_grid(int W, int H, int T) { return dim3((unsigned)((size_t)((W + 63) / 64) * H * T)); }

hipError_t launch_warp(const float *f2, const float *f1, int f1_per_frame, int T, int C, int H,
                       int W, const float *u, const float *v, size_t uv_frame_stride,
                       double h_col, double h_row, int acc64, float *out, hipStream_t st)
{
    if ((size_t)((W + 63) / 64) * H * T > 0x7fffffffu) return hipErrorInvalidValue;
    const dim3 g = warp_grid(W, H, T), b(64);
    if (acc64)
        hipLaunchKernelGGL((k_warp<float, float, double>), g, b, 0, st, f2, f1, f1_per_frame, C, H, W, u, v,
                           uv_frame_stride, h_col, h_row, out);
    else
        hipLaunchKernelGGL((k_warp<float, float, float>), g, b, 0, st, f2, f1, f1_per_frame, C, H, W, u, v,
                           uv_frame_stride, h_col, h_row, out);
    return hipGetLastError();
}

hipError_t launch_warp_ex(const void *f2, const void *f1, int img_f64, int acc64, int f1_per_frame, int T,
                          int C, int H, int W, const double *u, const double *v, size_t uv_frame_stride,
                          float *out, hipStream_t st)
{
    if ((size_t)((W + 63) / 64) * H * T > 0x7fffffffu) return hipErrorInvalidValue;
    const dim3 g = warp_grid(W, H, T), b(64);
    if (img_f64)
        hipLaunchKernelGGL((k_warp<double, double, double>), g, b, 0, st, (const double *)f2,
                           (const double *)f1, f1_per_frame, C, H, W, u, v, uv_frame_stride, 1.0, 1.0, out);
    else if (acc64)
        hipLaunchKernelGGL((k_warp<double, float, double>), g, b, 0, st, (const float *)f2, (const float *)f1,
                           f1_per_frame, C, H, W, u, v, uv_frame_stride, 1.0, 1.0, out);
    else
        hipLaunchKernelGGL((k_warp<double, float, float>), g, b, 0, st, (const float *)f2, (const float *)f1,
                           f1_per_frame, C, H, W, u, v, uv_frame_stride, 1.0, 1.0, out);
    return hipGetLastError();
}

// imregister_wrapper(..., interpolation_method="linear"): cv2.remap with
// INTER_LINEAR + BORDER_REPLICATE.  The map handling is that of k_warp (the
// float32 map from a float64 sum, out-of-bounds test on the unclipped map,
// clip, 1/32-px rounding); the tap origin is (X >> 5, Y >> 5) and the 2x2
// weights are wy[i] * wx[j] of the float table {1 - k/32, k/32}.  Origins
// with sx < W - 1 and sy < H - 1 read their 2x2 taps directly, the others
// clamp them to the image; both sum the four products left to right in Acc
// (float for float32 images, double for float64 images).
template <typename U, typename I, typename Acc>
__global__ void k_warp_linear(const I *__restrict__ f2, const I *__restrict__ f1, int f1_per_frame, int C,
                              int H, int W, const U *__restrict__ u, const U *__restrict__ v,
                              size_t uv_stride, float *__restrict__ out, int sat = 0, float lo = 0.f,
                              float hi = 0.f)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    if (x >= W) return;
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    const I *F2 = f2 + (size_t)t * C * HW;
    const I *F1 = f1 + (f1_per_frame ? (size_t)t * C * HW : 0);
    float *O = out + (size_t)t * C * HW;
    const float mx = (float)((double)x + (double)u[t * uv_stride + p]);
    const float my = (float)((double)y + (double)v[t * uv_stride + p]);
    if ((mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H)) {
        for (int c = 0; c < C; ++c) O[c * HW + p] = (float)F1[c * HW + p];
        return;
    }
    const float cx = mx > (float)(W - 1) ? (float)(W - 1) : mx;
    const float cy = my > (float)(H - 1) ? (float)(H - 1) : my;
    const int X = (int)rintf(cx * 32.f), Y = (int)rintf(cy * 32.f);
    const int sx = X >> 5, sy = Y >> 5;
    const float fx = (float)(X & 31) * (1.f / 32), fy = (float)(Y & 31) * (1.f / 32);
    const float wx0 = 1.f - fx, wy0 = 1.f - fy;
    const float w0 = wy0 * wx0, w1 = wy0 * fx, w2 = fy * wx0, w3 = fy * fx;
    const int x0 = clampi(sx, 0, W - 1), x1 = clampi(sx + 1, 0, W - 1);
    const int y0 = clampi(sy, 0, H - 1), y1 = clampi(sy + 1, 0, H - 1);
    for (int c = 0; c < C; ++c) {
        const I *S = F2 + c * HW;
        const Acc sum = (Acc)S[(size_t)y0 * W + x0] * (Acc)w0 + (Acc)S[(size_t)y0 * W + x1] * (Acc)w1 +
                        (Acc)S[(size_t)y1 * W + x0] * (Acc)w2 + (Acc)S[(size_t)y1 * W + x1] * (Acc)w3;
        O[c * HW + p] = warp_store_value((float)sum, sat, lo, hi);
    }
}

hipError_t launch_warp_linear_ex(const void *f2, const void *f1, int img_f64, int acc64, int f1_per_frame,
                                 int T, int C, int H, int W, const double *u, const double *v,
                                 size_t uv_frame_stride, float *out, hipStream_t st)
{
    const dim3 g((W + 63) / 64, H, T), b(64);
    if (img_f64)
        hipLaunchKernelGGL((k_warp_linear<double, double, double>), g, b, 0, st, (const double *)f2,
                           (const double *)f1, f1_per_frame, C, H, W, u, v, uv_frame_stride, out);
    else if (acc64)
        hipLaunchKernelGGL((k_warp_linear<double, float, double>), g, b, 0, st, (const float *)f2,
                           (const float *)f1, f1_per_frame, C, H, W, u, v, uv_frame_stride, out);
    else
        hipLaunchKernelGGL((k_warp_linear<double, float, float>), g, b, 0, st, (const float *)f2,
                           (const float *)f1, f1_per_frame, C, H, W, u, v, uv_frame_stride, out);
    return hipGetLastError();
}

hipError_t launch_warp_saturate(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C,
                                int H, int W, const double *u, const double *v, size_t uv_frame_stride,
                                float lo, float hi, float *out, hipStream_t st)
{
    const dim3 g((W + 63) / 64, H, T), b(64);
    if (linear)
        hipLaunchKernelGGL((k_warp_linear<double, float, float>), g, b, 0, st, f2, f1, f1_per_frame, C, H, W, u,
                           v, uv_frame_stride, out, 1, lo, hi);
    else if ((size_t)((W + 63) / 64) * H * T > 0x7fffffffu)
        return hipErrorInvalidValue;
    else
        hipLaunchKernelGGL((k_warp<double, float, float>), warp_grid(W, H, T), b, 0, st, f2, f1, f1_per_frame, C, H,
                           W, u, v, uv_frame_stride, 1.0, 1.0, out, 1, lo, hi);
    return hipGetLastError();
}

// cv2.remap on 8-bit images, OpenCV's 8U path (imgwarp.cpp: initInterTab2D,
// remapBicubic / remapBilinear with FixedPtCast<int, uchar, 15>).  The map
// handling is that of k_warp / k_warp_linear.  The weights of table cell
// (X & 31, Y & 31) are the float products wy[k1] * wx[k2] of the 1-D float
// table scaled by 2^15 and rounded to int16 (round half to even, saturated:
// a weight of exactly 1 becomes 32767); when they do not sum to 2^15 the
// difference is taken off the largest (sum too small) or put on the
// smallest (sum too large) cell of the window (k1, k2) in [KS/2, KS/2 + 2)^2,
// scanned row by row, first extreme kept.  For the 2x2 bilinear table the
// window reaches past the block, where OpenCV reads the next, still unfilled
// table cell: zeros here.  Each sample is the exact integer sum of its taps
// (clamped indices: BORDER_REPLICATE), stored as clamp((sum + 2^14) >> 15,
// 0, 255).  Samples arrive as float32 holding the uint8 values (exact).
template <bool LINEAR>
__global__ void k_warp_u8(const float *__restrict__ f2, const float *__restrict__ f1, int f1_per_frame, int C,
                          int H, int W, const double *__restrict__ u, const double *__restrict__ v,
                          size_t uv_stride, float *__restrict__ out)
{
    constexpr int KS = LINEAR ? 2 : 4, H2 = KS / 2;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    if (x >= W) return;
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    const float *F2 = f2 + (size_t)t * C * HW;
    const float *F1 = f1 + (f1_per_frame ? (size_t)t * C * HW : 0);
    float *O = out + (size_t)t * C * HW;
    const float mx = (float)((double)x + u[t * uv_stride + p]);
    const float my = (float)((double)y + v[t * uv_stride + p]);
    if ((mx < 0.f) || (mx >= (float)W) || (my < 0.f) || (my >= (float)H)) {
        for (int c = 0; c < C; ++c) O[c * HW + p] = F1[c * HW + p];
        return;
    }
    const float cx = mx > (float)(W - 1) ? (float)(W - 1) : mx;
    const float cy = my > (float)(H - 1) ? (float)(H - 1) : my;
    const int X = (int)rintf(cx * 32.f), Y = (int)rintf(cy * 32.f);
    const int sx = (X >> 5) - (LINEAR ? 0 : 1), sy = (Y >> 5) - (LINEAR ? 0 : 1);
    float tx[4], ty[4];
    if (LINEAR) {
        tx[1] = (float)(X & 31) * (1.f / 32); tx[0] = 1.f - tx[1];
        ty[1] = (float)(Y & 31) * (1.f / 32); ty[0] = 1.f - ty[1];
    } else {
        cv_cubic_coeffs((float)(X & 31) * (1.f / 32), tx);
        cv_cubic_coeffs((float)(Y & 31) * (1.f / 32), ty);
    }
    int wt[KS * KS];
    int isum = 0;
#pragma unroll
    for (int k1 = 0; k1 < KS; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < KS; ++k2) {
            const float wv = ty[k1] * tx[k2];
            const int r = clampi((int)rintf(wv * 32768.f), -32768, 32767);
            wt[k1 * KS + k2] = r;
            isum += r;
        }
    if (isum != 32768) {
        // window cells past the KS x KS block read as 0
        int mi = H2 * KS + H2, Mi = mi, mval = wt[mi], Mval = wt[Mi];
#pragma unroll
        for (int k1 = H2; k1 < H2 + 2; ++k1)
#pragma unroll
            for (int k2 = H2; k2 < H2 + 2; ++k2) {
                const int k = k1 * KS + k2;
                const int val = (k1 < KS && k2 < KS) ? wt[(k1 < KS && k2 < KS) ? k : 0] : 0;
                if (val < mval) { mi = k; mval = val; }
                else if (val > Mval) { Mi = k; Mval = val; }
            }
        const int diff = isum - 32768, tgt = diff < 0 ? Mi : mi;
#pragma unroll
        for (int k = 0; k < KS * KS; ++k)
            if (k == tgt) wt[k] = (int)(short)(wt[k] - diff);
    }
    int xs[KS], ys[KS];
#pragma unroll
    for (int i = 0; i < KS; ++i) {
        xs[i] = clampi(sx + i, 0, W - 1);
        ys[i] = clampi(sy + i, 0, H - 1);
    }
    for (int c = 0; c < C; ++c) {
        const float *S = F2 + c * HW;
        int sum = 0;
#pragma unroll
        for (int i = 0; i < KS; ++i)
#pragma unroll
            for (int j = 0; j < KS; ++j) sum += (int)S[(size_t)ys[i] * W + xs[j]] * wt[i * KS + j];
        O[c * HW + p] = (float)clampi((sum + (1 << 14)) >> 15, 0, 255);
    }
}

hipError_t launch_warp_u8(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C, int H,
                          int W, const double *u, const double *v, size_t uv_frame_stride, float *out,
                          hipStream_t st)
{
    const dim3 g((W + 63) / 64, H, T), b(64);
    if (linear)
        hipLaunchKernelGGL(k_warp_u8<true>, g, b, 0, st, f2, f1, f1_per_frame, C, H, W, u, v, uv_frame_stride,
                           out);
    else
        hipLaunchKernelGGL(k_warp_u8<false>, g, b, 0, st, f2, f1, f1_per_frame, C, H, W, u, v, uv_frame_stride,
                           out);
    return hipGetLastError();
}

// numpy's np.mean(w, axis=0) of a C-contiguous float32 (n, len) array: the
// reduction over the outer axis adds the rows one after another in float32
// (pairwise summation applies only along a contiguous inner axis), and the
// float32 sum divided by the count rounds as a float32 division would (the
// quotient computed in float64 and rounded to float32 is the correctly
// rounded float32 quotient, since 53 >= 2 * 24 + 2).
__global__ void k_mean_frames_f32(const float *__restrict__ w, int n, size_t len, float *__restrict__ out)
{
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= len) return;
    float s = w[j];
    for (int i = 1; i < n; ++i) s += w[(size_t)i * len + j];
    out[j] = s / (float)n;
}

hipError_t launch_mean_frames_f32(const float *w, int n, size_t len, float *out, hipStream_t st)
{
    const size_t nb = (len + 255) / 256;
    if (nb == 0 || nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_mean_frames_f32, dim3((unsigned)nb), dim3(256), 0, st, w, n, len, out);
    return hipGetLastError();
}

// ---------------------------------------------------------- motion tensor
// core/optical_flow.py:73-345 (get_motion_tensor_gc / _gray / _cs), one
// cell of one channel.  Symmetric 1-px pads are clamped reads; np.gradient
// is central in the interior; the re-padded ring of the first derivatives
// copies the nearest interior value.
//
// numpy evaluates each expression in the dtype of its array operands: the
// level image f1 is float64 (the resize casts back to the input dtype);
// the warped image f2 is float32 below the top level (imregister_wrapper
// returns float32) and float64 at the top level.  P1 / P2 mark a float32
// input and PM = P1 && P2 the terms that combine both.  A float32 operation
// is evaluated as the double operation rounded to float32 (exact for
// + - * / sqrt of float32 operands, since double has more than 2 * 24 + 2
// bits); a Python scalar meets a float32 array as a float32.
enum TensorKind { kTensorGC = 0, kTensorGray = 1, kTensorCS = 2 };

struct TensorOut {
    double J11, J22, J33, J12, J13, J23;
};

template <bool F32>
__device__ inline double rnd(double x)
{
    if constexpr (F32)
        return (double)(float)x;
    else
        return x;
}

// np.gradient interior value (x[+1] - x[-1]) / (2. * h).  UNIT: h == 1 (the
// finest level), where dividing by 2 is the exact scaling by 0.5.
template <bool F32, bool UNIT = false>
__device__ inline double dgrad(double xp, double xm, double h)
{
    if constexpr (UNIT)
        return rnd<F32>(rnd<F32>(xp - xm) * 0.5);
    else
        return rnd<F32>(rnd<F32>(xp - xm) / rnd<F32>(2.0 * h));
}

// (x[-1] - 2 * x[0] + x[+1]) / (h ** 2); UNIT: h == 1, the division is exact
template <bool F32, bool UNIT = false>
__device__ inline double dsecond(double xm, double x0, double xp, double h)
{
    if constexpr (UNIT)
        return rnd<F32>(rnd<F32>(xm - 2 * x0) + xp);
    else
        return rnd<F32>(rnd<F32>(rnd<F32>(xm - 2 * x0) + xp) / rnd<F32>(h * h));
}

__device__ inline double img(const float *f, int w, int r, int c) { return (double)f[(size_t)r * w + c]; }

struct Clamp3 {
    int ru, rd, cl, cr;
};

__device__ inline Clamp3 clamp3(int h, int w, int r, int c)
{
    Clamp3 k;
    k.ru = r > 0 ? r - 1 : 0;
    k.rd = r < h - 1 ? r + 1 : h - 1;
    k.cl = c > 0 ? c - 1 : 0;
    k.cr = c < w - 1 ? c + 1 : w - 1;
    return k;
}

template <bool P1, bool P2, bool UNIT = false>
__device__ inline TensorOut tensor_gc(const float *f1, const float *f2, int h, int w, int r, int c,
                                      double hy, double hx)
{
    constexpr bool PM = P1 && P2;
    const Clamp3 k = clamp3(h, w, r, c);
    // fx = 0.5 (d/dx f1 + d/dx f2) on rows r-1, r+1 (clamped: the re-pad)
    auto fx_at = [&](int rr) {
        double g1 = dgrad<P1, UNIT>(img(f1, w, rr, k.cr), img(f1, w, rr, k.cl), hx);
        double g2 = dgrad<P2, UNIT>(img(f2, w, rr, k.cr), img(f2, w, rr, k.cl), hx);
        return rnd<PM>(0.5 * rnd<PM>(g1 + g2));
    };
    auto ft_at = [&](int rr, int cc) { return rnd<PM>(img(f2, w, rr, cc) - img(f1, w, rr, cc)); };
    const double fxy = dgrad<PM, UNIT>(fx_at(k.rd), fx_at(k.ru), hy);
    const double fxt = dgrad<PM, UNIT>(ft_at(r, k.cr), ft_at(r, k.cl), hx);
    const double fyt = dgrad<PM, UNIT>(ft_at(k.rd, c), ft_at(k.ru, c), hy);
    const double a0 = img(f1, w, r, c), b0 = img(f2, w, r, c);
    const double fxx1 = dsecond<P1, UNIT>(img(f1, w, r, k.cl), a0, img(f1, w, r, k.cr), hx);
    const double fyy1 = dsecond<P1, UNIT>(img(f1, w, k.ru, c), a0, img(f1, w, k.rd, c), hy);
    const double fxx2 = dsecond<P2, UNIT>(img(f2, w, r, k.cl), b0, img(f2, w, r, k.cr), hx);
    const double fyy2 = dsecond<P2, UNIT>(img(f2, w, k.ru, c), b0, img(f2, w, k.rd, c), hy);
    const double fxx = rnd<PM>(0.5 * rnd<PM>(fxx1 + fxx2));
    const double fyy = rnd<PM>(0.5 * rnd<PM>(fyy1 + fyy2));
    // reg = 1.0 / ((np.sqrt(a**2 + b**2) ** 2) + 1e-6)
    const double sx = rnd<PM>(sqrt(rnd<PM>(rnd<PM>(fxx * fxx) + rnd<PM>(fxy * fxy))));
    const double sy = rnd<PM>(sqrt(rnd<PM>(rnd<PM>(fxy * fxy) + rnd<PM>(fyy * fyy))));
    const double regx = rnd<PM>(1.0 / rnd<PM>(rnd<PM>(sx * sx) + rnd<PM>(1e-6)));
    const double regy = rnd<PM>(1.0 / rnd<PM>(rnd<PM>(sy * sy) + rnd<PM>(1e-6)));
    // reg_x * fxx**2 is reg_x * (fxx * fxx); products left to right
    TensorOut o;
    o.J11 = rnd<PM>(rnd<PM>(regx * rnd<PM>(fxx * fxx)) + rnd<PM>(regy * rnd<PM>(fxy * fxy)));
    o.J22 = rnd<PM>(rnd<PM>(regx * rnd<PM>(fxy * fxy)) + rnd<PM>(regy * rnd<PM>(fyy * fyy)));
    o.J33 = rnd<PM>(rnd<PM>(regx * rnd<PM>(fxt * fxt)) + rnd<PM>(regy * rnd<PM>(fyt * fyt)));
    o.J12 = rnd<PM>(rnd<PM>(rnd<PM>(regx * fxx) * fxy) + rnd<PM>(rnd<PM>(regy * fxy) * fyy));
    o.J13 = rnd<PM>(rnd<PM>(rnd<PM>(regx * fxx) * fxt) + rnd<PM>(rnd<PM>(regy * fxy) * fyt));
    o.J23 = rnd<PM>(rnd<PM>(rnd<PM>(regx * fxy) * fxt) + rnd<PM>(rnd<PM>(regy * fyy) * fyt));
    return o;
}

// core/optical_flow.py:163-228
template <bool P1, bool P2, bool UNIT = false>
__device__ inline TensorOut tensor_gray(const float *f1, const float *f2, int h, int w, int r,
                                        int c, double hy, double hx)
{
    constexpr bool PM = P1 && P2;
    const Clamp3 k = clamp3(h, w, r, c);
    const double fx1 = dgrad<P1, UNIT>(img(f1, w, r, k.cr), img(f1, w, r, k.cl), hx);
    const double fy1 = dgrad<P1, UNIT>(img(f1, w, k.rd, c), img(f1, w, k.ru, c), hy);
    const double fx2 = dgrad<P2, UNIT>(img(f2, w, r, k.cr), img(f2, w, r, k.cl), hx);
    const double fy2 = dgrad<P2, UNIT>(img(f2, w, k.rd, c), img(f2, w, k.ru, c), hy);
    const double fx = rnd<PM>(0.5 * rnd<PM>(fx1 + fx2));
    const double fy = rnd<PM>(0.5 * rnd<PM>(fy1 + fy2));
    const double ft = rnd<PM>(img(f2, w, r, c) - img(f1, w, r, c));
    TensorOut o;
    o.J11 = rnd<PM>(fx * fx);
    o.J22 = rnd<PM>(fy * fy);
    o.J33 = rnd<PM>(ft * ft);
    o.J12 = rnd<PM>(fx * fy);
    o.J13 = rnd<PM>(fx * ft);
    o.J23 = rnd<PM>(fy * ft);
    return o;
}

// smoothed Heaviside 0.5 * (1.0 + r / np.sqrt(r * r + eps2)) of the
// directional difference r = (neighbour - centre) / dist
template <bool F32>
__device__ inline double census_s(double nb, double ce, double dist, double eps2)
{
    // a unit distance (axis neighbours at unit spacing) divides exactly
    const double rr = dist == 1.0 ? rnd<F32>(nb - ce) : rnd<F32>(rnd<F32>(nb - ce) / rnd<F32>(dist));
    const double d = rnd<F32>(sqrt(rnd<F32>(rnd<F32>(rr * rr) + rnd<F32>(eps2))));
    return rnd<F32>(0.5 * rnd<F32>(1.0 + rnd<F32>(rr / d)));
}

// core/optical_flow.py:231-345: the eight neighbour offsets in row-major
// order; per offset, the census images of both frames on the h x w core
// (edge-padded, i.e. clamped) and their np.gradient derivatives at the
// cell; float64 accumulation, then the scale 2 / 8.
template <bool P1, bool P2, bool UNIT = false>
__device__ inline TensorOut tensor_cs(const float *f1, const float *f2, int h, int w, int r, int c,
                                      double hy, double hx, double eps)
{
    constexpr bool PM = P1 && P2;
    const double eps2 = eps * eps;
    const Clamp3 k = clamp3(h, w, r, c);
    double A0 = 0, A1 = 0, A2 = 0, A3 = 0, A4 = 0, A5 = 0;
    for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            if (dy == 0 && dx == 0) continue;
            const double dist = sqrt((hy * dy) * (hy * dy) + (hx * dx) * (hx * dx));
            // census value of image f at core pixel (rr, cc) for this offset;
            // the neighbour comes from the symmetric pad (clamped)
            auto s_at = [&](const float *f, bool f32, int rr, int cc) {
                int nr = rr + dy, nc = cc + dx;
                nr = nr < 0 ? 0 : (nr > h - 1 ? h - 1 : nr);
                nc = nc < 0 ? 0 : (nc > w - 1 ? w - 1 : nc);
                const double nb = img(f, w, nr, nc), ce = img(f, w, rr, cc);
                return f32 ? census_s<true>(nb, ce, dist, eps2) : census_s<false>(nb, ce, dist, eps2);
            };
            const double sx1 = dgrad<P1, UNIT>(s_at(f1, P1, r, k.cr), s_at(f1, P1, r, k.cl), hx);
            const double sx2 = dgrad<P2, UNIT>(s_at(f2, P2, r, k.cr), s_at(f2, P2, r, k.cl), hx);
            const double sy1 = dgrad<P1, UNIT>(s_at(f1, P1, k.rd, c), s_at(f1, P1, k.ru, c), hy);
            const double sy2 = dgrad<P2, UNIT>(s_at(f2, P2, k.rd, c), s_at(f2, P2, k.ru, c), hy);
            const double sx = rnd<PM>(0.5 * rnd<PM>(sx1 + sx2));
            const double sy = rnd<PM>(0.5 * rnd<PM>(sy1 + sy2));
            const double st = rnd<PM>(s_at(f2, P2, r, c) - s_at(f1, P1, r, c));
            A0 = A0 + rnd<PM>(sx * sx);
            A1 = A1 + rnd<PM>(sy * sy);
            A2 = A2 + rnd<PM>(st * st);
            A3 = A3 + rnd<PM>(sx * sy);
            A4 = A4 + rnd<PM>(sx * st);
            A5 = A5 + rnd<PM>(sy * st);
        }
    const double scale = 2.0 / 8.0;
    TensorOut o;
    o.J11 = A0 * scale;
    o.J22 = A1 * scale;
    o.J33 = A2 * scale;
    o.J12 = A3 * scale;
    o.J13 = A4 * scale;
    o.J23 = A5 * scale;
    return o;
}

template <int KIND, bool P1, bool P2, bool UNIT = false>
__device__ inline TensorOut tensor_cell(const float *f1, const float *f2, int h, int w, int r, int c,
                                        double hy, double hx, double eps)
{
    if constexpr (KIND == kTensorGC)
        return tensor_gc<P1, P2, UNIT>(f1, f2, h, w, r, c, hy, hx);
    else if constexpr (KIND == kTensorGray)
        return tensor_gray<P1, P2, UNIT>(f1, f2, h, w, r, c, hy, hx);
    else
        return tensor_cs<P1, P2, UNIT>(f1, f2, h, w, r, c, hy, hx, eps);
}

// cu = sum_nb w_nb (u_nb - u_c) over (left, right, down, up); u is
// edge-padded, so neighbours outside the level contribute exactly 0.
__device__ inline double smooth_rhs(const float *u, int h, int w, int r, int c, double ax,
                                    double ay)
{
    double uc = img(u, w, r, c);
    double ul = img(u, w, r, c > 0 ? c - 1 : c), ur = img(u, w, r, c < w - 1 ? c + 1 : c);
    double ud = img(u, w, r < h - 1 ? r + 1 : r, c), uu = img(u, w, r > 0 ? r - 1 : r, c);
    double s = ax * (ul - uc);
    s = s + ax * (ur - uc);
    s = s + ay * (ud - uc);
    s = s + ay * (uu - uc);
    return s;
}

// Tiles of kTensorRows rows x kTensorCols columns, one thread per cell
// (lanes run along a row, so the image reads are coalesced).  Tiles are
// numbered frame-major, then row-major, and dealt to the XCDs in contiguous
// ranges (xcd_block) so neighbouring tiles share one L2.  J planes
// (component-major, then channel) and cu, cv are written skewed: the output
// planes of a tile go through LDS (a channel's six at once) and out diagonal
// by diagonal, so that 16 consecutive lanes store one diagonal's 16 rows --
// a contiguous 128-byte run -- instead of every lane of a row storing to a
// different diagonal.  512-thread work-groups (24 KB of LDS) let several
// tiles share a CU, so one tile's staging barriers do not stall the CU.  The level image
// f1 is float64 data in the reference; P2 marks a float32 warped image
// (every level below the top).  UNIT: unit grid spacing (hy = hx = 1, the
// finest level), where the gradients' divisions by 2h and h^2 are exact.
constexpr int kTensorRows = 16, kTensorCols = 32;
constexpr int kTensorDiags = kTensorRows + kTensorCols - 1;  // diagonals a tile touches

// One tile (rows r0.., columns c0..) of one frame: f1t, f2t its C level
// planes, ut, vt its level flow, Jt, ct its skewed output planes.
template <int KIND, bool P2, bool UNIT = false>
__device__ __forceinline__ void tensor_tile(double (*stage)[kTensorRows][kTensorCols], const float *__restrict__ f1t,
                                            const float *__restrict__ f2t, int C, int h, int w, int hs, size_t plane,
                                            double hy, double hx, double eps, const float *__restrict__ ut,
                                            const float *__restrict__ vt, double ax, double ay, double *__restrict__ Jt,
                                            double *__restrict__ ct, int r0, int c0)
{
    const int r = r0 + threadIdx.y, c = c0 + threadIdx.x;
    const bool ok = r < h && c < w;  // every thread takes part in the staging barriers
    const int tid = threadIdx.y * kTensorCols + threadIdx.x;
    const size_t hw = (size_t)h * w, d0 = (size_t)(r0 + c0);
    // N output planes of the tile at once: cell (r, c) -> slot (r + c) * hs + r,
    // stored diagonal by diagonal (slot s = diagonal j of the tile, row i);
    // plane e of the N goes to base + e * step
    auto put = [&](auto n_planes, const double *vals, double *base, size_t step) {
        constexpr int N = decltype(n_planes)::value;
#pragma unroll
        for (int e = 0; e < N; ++e) stage[e][threadIdx.y][threadIdx.x] = vals[e];
        __syncthreads();
        for (int s = tid; s < kTensorDiags * kTensorRows; s += kTensorCols * kTensorRows) {
            const int j = s / kTensorRows, i = s % kTensorRows, cl = j - i;
            if (cl >= 0 && cl < kTensorCols && r0 + i < h && c0 + cl < w) {
                const size_t q = (d0 + j) * hs + r0 + i;
#pragma unroll
                for (int e = 0; e < N; ++e) base[e * step + q] = stage[e][i][cl];
            }
        }
        __syncthreads();
    };
    for (int ch = 0; ch < C; ++ch) {
        TensorOut o{};
        if (ok)
            o = tensor_cell<KIND, false, P2, UNIT>(f1t + ch * hw, f2t + ch * hw, h, w, r, c, hy, hx, eps);
        // components J11, J22, J33, J12, J13, J23 -> planes (e * C + ch)
        const double vals[6] = {o.J11, o.J22, o.J33, o.J12, o.J13, o.J23};
        put(std::integral_constant<int, 6>{}, vals, Jt + (size_t)ch * plane, (size_t)C * plane);
    }
    const double cvals[2] = {ok ? smooth_rhs(ut, h, w, r, c, ax, ay) : 0.0,
                             ok ? smooth_rhs(vt, h, w, r, c, ax, ay) : 0.0};
    put(std::integral_constant<int, 2>{}, cvals, ct, plane);
}

template <int KIND, bool P2, bool UNIT = false>
__global__ __launch_bounds__(kTensorCols * kTensorRows, KIND == kTensorCS ? 2 : 6) void k_tensor(
    const float *__restrict__ f1, const float *__restrict__ f2, int C, int h, int w, int hs,
    size_t plane, double hy, double hx, double eps, size_t f1_fstride, const float *__restrict__ u,
    const float *__restrict__ v, size_t uv_stride, double ax, double ay, double *__restrict__ J,
    double *__restrict__ cuv, size_t fstride, int T)
{
    __shared__ double stage[6][kTensorRows][kTensorCols];
    const int tiles_x = (w + kTensorCols - 1) / kTensorCols, tiles_y = (h + kTensorRows - 1) / kTensorRows;
    const unsigned nb = (unsigned)tiles_x * tiles_y * T;
    const unsigned L = xcd_block(blockIdx.x, nb);
    const int t = L / (tiles_x * tiles_y), rem = L % (tiles_x * tiles_y);
    const int r0 = (rem / tiles_x) * kTensorRows, c0 = (rem % tiles_x) * kTensorCols;
    tensor_tile<KIND, P2, UNIT>(stage, f1 + (size_t)t * f1_fstride, f2 + (size_t)t * C * h * w, C, h, w, hs, plane,
                                hy, hx, eps, u + t * uv_stride, v + t * uv_stride, ax, ay, J + (size_t)t * fstride,
                                cuv + (size_t)t * fstride, r0, c0);
}

hipError_t launch_tensor(int kind, int f2_f32, const float *f1, size_t f1_frame_stride, const float *f2,
                         int T, int C, int h, int w, double hy, double hx, double eps, const float *u,
                         const float *v, size_t uv_frame_stride, double ax, double ay, double *J,
                         double *cuv, size_t frame_stride, const Skew &s, hipStream_t st)
{
    const size_t nb = (size_t)((w + kTensorCols - 1) / kTensorCols) * ((h + kTensorRows - 1) / kTensorRows) * T;
    if (nb > 0x7fffffffu) return hipErrorInvalidValue;
#define FRH_TEN(K, P)                                                                             \
    hipLaunchKernelGGL((k_tensor<K, P>), dim3((unsigned)nb), dim3(kTensorCols, kTensorRows), 0, st, f1, f2, C, \
                       h, w, s.hs, s.plane, hy, hx, eps, f1_frame_stride, u, v, uv_frame_stride, ax, \
                       ay, J, cuv, frame_stride, T)
#define FRH_TEN_U(K, P)                                                                           \
    hipLaunchKernelGGL((k_tensor<K, P, true>), dim3((unsigned)nb), dim3(kTensorCols, kTensorRows), 0, st, f1, f2, \
                       C, h, w, s.hs, s.plane, hy, hx, eps, f1_frame_stride, u, v, uv_frame_stride, ax, \
                       ay, J, cuv, frame_stride, T)
    const bool unit = hy == 1.0 && hx == 1.0;
    switch (kind * 2 + (f2_f32 ? 1 : 0)) {
    case 0:
        if (unit) FRH_TEN_U(kTensorGC, false);
        else FRH_TEN(kTensorGC, false);
        break;
    case 1:
        if (unit) FRH_TEN_U(kTensorGC, true);
        else FRH_TEN(kTensorGC, true);
        break;
    case 2:
        if (unit) FRH_TEN_U(kTensorGray, false);
        else FRH_TEN(kTensorGray, false);
        break;
    case 3:
        if (unit) FRH_TEN_U(kTensorGray, true);
        else FRH_TEN(kTensorGray, true);
        break;
    case 4:
        if (unit) FRH_TEN_U(kTensorCS, false);
        else FRH_TEN(kTensorCS, false);
        break;
    case 5:
        if (unit) FRH_TEN_U(kTensorCS, true);
        else FRH_TEN(kTensorCS, true);
        break;
    default: return hipErrorInvalidValue;
    }
#undef FRH_TEN
#undef FRH_TEN_U
    return hipGetLastError();
}

// Row-major tensor for the stand-alone motion-tensor entry point:
// (h+2, w+2) outputs with a zero ring.
template <int KIND, bool P1, bool P2>
__global__ void k_tensor_rowmajor(const float *__restrict__ f1, const float *__restrict__ f2,
                                  int h, int w, double hy, double hx, double eps, double *J11,
                                  double *J22, double *J33, double *J12, double *J13, double *J23)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    int wp = w + 2;
    if (i >= wp) return;
    size_t q = (size_t)j * wp + i;
    if (j == 0 || j == h + 1 || i == 0 || i == w + 1) {
        J11[q] = J22[q] = J33[q] = J12[q] = J13[q] = J23[q] = 0.0;
        return;
    }
    TensorOut o = tensor_cell<KIND, P1, P2>(f1, f2, h, w, j - 1, i - 1, hy, hx, eps);
    J11[q] = o.J11;
    J22[q] = o.J22;
    J33[q] = o.J33;
    J12[q] = o.J12;
    J13[q] = o.J13;
    J23[q] = o.J23;
}

hipError_t launch_tensor_rowmajor(int kind, int f1_f32, int f2_f32, const float *f1, const float *f2,
                                  int h, int w, double hy, double hx, double eps, double *J11,
                                  double *J22, double *J33, double *J12, double *J13, double *J23,
                                  hipStream_t st)
{
    dim3 g((w + 2 + 63) / 64, h + 2), b(64);
#define FRH_TRM(K, A, B)                                                                          \
    hipLaunchKernelGGL((k_tensor_rowmajor<K, A, B>), g, b, 0, st, f1, f2, h, w, hy, hx, eps, J11, J22, \
                       J33, J12, J13, J23)
    switch (kind * 4 + (f1_f32 ? 2 : 0) + (f2_f32 ? 1 : 0)) {
    case 0: FRH_TRM(kTensorGC, false, false); break;
    case 1: FRH_TRM(kTensorGC, false, true); break;
    case 2: FRH_TRM(kTensorGC, true, false); break;
    case 3: FRH_TRM(kTensorGC, true, true); break;
    case 4: FRH_TRM(kTensorGray, false, false); break;
    case 5: FRH_TRM(kTensorGray, false, true); break;
    case 6: FRH_TRM(kTensorGray, true, false); break;
    case 7: FRH_TRM(kTensorGray, true, true); break;
    case 8: FRH_TRM(kTensorCS, false, false); break;
    case 9: FRH_TRM(kTensorCS, false, true); break;
    case 10: FRH_TRM(kTensorCS, true, false); break;
    case 11: FRH_TRM(kTensorCS, true, true); break;
    default: return hipErrorInvalidValue;
    }
#undef FRH_TRM
    return hipGetLastError();
}

// ------------------------------------------------------ coefficient assembly
// Lag-block coefficients of core/level_solver.py:229-255 + :262-351 for one
// cell (ordering documented with the oracle's assemble_pixel):
//   Du = (((wl+wr)+wd)+wu) + sum_k Wpsi_k J11_k,  Dv likewise with J22,
//   A12 = sum_k Wpsi_k J12_k, Bu = cu - sum_k Wpsi_k J13_k, Bv = cv - ...J23,
//   stored as {A12, Bu, Bv, omega/Du, omega/Dv}.

// Coefficients of one cell from its raw data: J (6 components x C channels,
// component-major), W (C), cu, cv, the sum ws of its four smoothness
// weights and the cell's start values (x, y).
template <int C>
__device__ __forceinline__ void coef_from_raw_ws(const double *J, const double *Wc, double cu,
                                                 double cv, double x, double y, double ws,
                                                 const AData &ad, int psi_one, double *K)
{
    double Du = ws, Dv = ws, A12 = 0.0, Bu = cu, Bv = cv;
#pragma unroll
    for (int k = 0; k < C; ++k) {
        double j11 = J[0 * C + k], j22 = J[1 * C + k], j33 = J[2 * C + k];
        double j12 = J[3 * C + k], j13 = J[4 * C + k], j23 = J[5 * C + k];
        double psi = 1.0;
        if (!psi_one) {
            double val = j11 * x * x + j22 * y * y + j23 * y + 2 * j12 * x * y + 2 * j13 * x +
                         j23 * y + j33;
            if (val < 0) val = 0;
            psi = ad.a[k] * frh_det_pow(val + 1e-5, ad.am1[k]);
            if (ad.gnc[k]) psi = ad.mix + ad.beta * psi;
        }
        double wp = Wc[k] * psi;
        Du = Du + wp * j11;
        Dv = Dv + wp * j22;
        A12 = A12 + wp * j12;
        Bu = Bu - wp * j13;
        Bv = Bv - wp * j23;
    }
    const double omega = 1.95;
    K[0] = A12;
    K[1] = Bu;
    K[2] = Bv;
    K[3] = Du != 0 ? omega / Du : 0.0;
    K[4] = Dv != 0 ? omega / Dv : 0.0;
}

// The same with the constant weights ax (left, right) and ay (down, up).
template <int C>
__device__ __forceinline__ void coef_from_raw(const double *J, const double *Wc, double cu,
                                              double cv, double x, double y, double ax, double ay,
                                              const AData &ad, int psi_one, double *K)
{
    coef_from_raw_ws<C>(J, Wc, cu, cv, x, y, ((ax + ax) + ay) + ay, ad, psi_one, K);
}

// WT: the weight planes' type (float32 in the pyramid solver, whose weight
// levels are float32 values; float64 for the stand-alone level solver's
// caller-given weights).
// Slot q of one frame: Jt, ct, dt, co its J, cu/cv, du/dv and coefficient
// planes; Wsk the level's weight planes.
template <int C, typename WT>
__device__ __forceinline__ void assemble_cell(const double *__restrict__ Jt, const WT *__restrict__ Wsk,
                                              const double *__restrict__ ct, const double *__restrict__ dt,
                                              double *__restrict__ co, size_t plane, size_t q, double ax,
                                              double ay, const AData &ad, int psi_one, int zero_init)
{
    // J, cu/cv, du/dv and the coefficients stream through once per pass
    // (gigabytes apart): non-temporal accesses; the weight planes, shared
    // by every frame, stay cacheable
    double x = 0.0, y = 0.0;
    if (!psi_one && !zero_init) {
        x = __builtin_nontemporal_load(&dt[q]);
        y = __builtin_nontemporal_load(&dt[plane + q]);
    }
    double Jr[6 * C], Wr[C], K[5];
#pragma unroll
    for (int e = 0; e < 6 * C; ++e) Jr[e] = __builtin_nontemporal_load(&Jt[(size_t)e * plane + q]);
#pragma unroll
    for (int k = 0; k < C; ++k) Wr[k] = (double)Wsk[(size_t)k * plane + q];
    coef_from_raw<C>(Jr, Wr, __builtin_nontemporal_load(&ct[q]), __builtin_nontemporal_load(&ct[plane + q]), x, y,
                     ax, ay, ad, psi_one, K);
#pragma unroll
    for (int e = 0; e < 5; ++e) __builtin_nontemporal_store(K[e], &co[e * plane + q]);
}

template <int C, typename WT>
__global__ void k_assemble(const double *__restrict__ J, const WT *__restrict__ Wsk,
                           const double *__restrict__ cuv, const double *__restrict__ duv,
                           double *__restrict__ coef, size_t fstride, int h, int w, int hs,
                           size_t plane, double ax, double ay, AData ad, int psi_one,
                           int zero_init)
{
    int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x, t = blockIdx.y;
    int c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    const size_t fo = (size_t)t * fstride;
    assemble_cell<C, WT>(J + fo, Wsk, cuv + fo, duv + fo, coef + fo, plane, (size_t)d * hs + r, ax, ay, ad,
                         psi_one, zero_init);
}

template <typename WT>
static hipError_t launch_assemble_t(const double *J, const WT *Wsk, const double *cuv, const double *duv,
                                    double *coef, size_t frame_stride, int T, int C, const Skew &s, double ax,
                                    double ay, const AData &ad, int psi_one, int zero_init, hipStream_t st)
{
    // one work-group per diagonal and frame (and per 1024 rows of taller levels)
    dim3 g(s.h + s.w - 1, T, (s.hs + 1023) / 1024), b(s.hs < 1024 ? s.hs : 1024);
#define FRH_ASM(CC)                                                                            \
    hipLaunchKernelGGL((k_assemble<CC, WT>), g, b, 0, st, J, Wsk, cuv, duv, coef, frame_stride, s.h, \
                       s.w, s.hs, s.plane, ax, ay, ad, psi_one, zero_init)
    switch (C) {
    case 1: FRH_ASM(1); break;
    case 2: FRH_ASM(2); break;
    case 3: FRH_ASM(3); break;
    case 4: FRH_ASM(4); break;
    case 5: FRH_ASM(5); break;
    case 6: FRH_ASM(6); break;
    case 7: FRH_ASM(7); break;
    case 8: FRH_ASM(8); break;
    default: return hipErrorInvalidValue;
    }
#undef FRH_ASM
    return hipGetLastError();
}

hipError_t launch_assemble(const double *J, const double *Wsk, const double *cuv,
                           const double *duv, double *coef, size_t frame_stride, int T, int C,
                           const Skew &s, double ax, double ay, const double *a_data_minus1,
                           const double *a_data, int psi_one, int zero_init, double gnc_beta,
                           hipStream_t st)
{
    const AData ad = make_adata(C, a_data, a_data_minus1, gnc_beta);
    return launch_assemble_t<double>(J, Wsk, cuv, duv, coef, frame_stride, T, C, s, ax, ay, ad, psi_one,
                                            zero_init, st);
}

hipError_t launch_assemble_w32(const double *J, const float *Wsk, const double *cuv, const double *duv,
                               double *coef, size_t frame_stride, int T, int C, const Skew &s, double ax,
                               double ay, const double *a_data_minus1, const double *a_data, int psi_one,
                               int zero_init, double gnc_beta, hipStream_t st)
{
    const AData ad = make_adata(C, a_data, a_data_minus1, gnc_beta);
    return launch_assemble_t<float>(J, Wsk, cuv, duv, coef, frame_stride, T, C, s, ax, ay, ad, psi_one,
                                    zero_init, st);
}

// -------------------------------- nonlinear smoothness (a_smooth != 1)
// core/level_solver.py:69-155,248-255,262-316.  The smoothness weight of
// the edge between a cell and its neighbour is 0.5 (psi_s[c] + psi_s[nb])
// alpha / h^2, with psi_s refreshed over the padded grid at every psi
// refresh, so each lag block carries nine coefficients per cell:
// {A12, Bu, Bv, omega/Du, omega/Dv, w_left, w_right, w_down, w_up}.

// The level's initial flow on the padded (h+2) x (w+2) grid: either the
// caller's padded float64 planes (u64, v64; one frame) or the level's
// float32 flow (u32, v32; frame stride fs32) read with the edge pad of
// add_boundary (clamped indices).
struct PadFlow {
    const float *u32, *v32;
    size_t fs32;
    const double *u64, *v64;
};

__device__ inline double pad_flow(const PadFlow &f, int comp, int t, int h, int w, int jp, int ip)
{
    if (f.u64) return (comp ? f.v64 : f.u64)[(size_t)jp * (w + 2) + ip];
    const int r = clampi(jp - 1, 0, h - 1), c = clampi(ip - 1, 0, w - 1);
    return (double)(comp ? f.v32 : f.u32)[(size_t)t * f.fs32 + (size_t)r * w + c];
}

// nonlinearity_smoothness_2d at a psi refresh: psi_s over the padded grid
// of u + du, v + dv (one thread per padded cell).  du, dv of the interior
// come from the skewed state; the padded ring holds what set_boundary_2d
// copied before the previous sweep -- the start values of that sweep at
// the nearest border cell, which the sweep recorded in the ring planes.
// zero = 1 before the first sweep (du = 0 everywhere).  One-sided
// differences on the padded edge, central inside; psi_s = a (t + 1e-5)^
// (a - 1) with frh_det_pow, then (1 - beta) + beta psi_s when mix is set
// (compute_flow_gnc).
// u + du (comp 0) or v + dv (comp 1) at padded cell (j, i) of frame t
__device__ __forceinline__ double psi_full(const PadFlow &fl, const double *D, const double *R, int comp, int t,
                                           int h, int w, int hs, size_t plane, int zero, int j, int i)
{
    double d = 0.0;
    if (!zero) {
        const bool inner = j >= 1 && j <= h && i >= 1 && i <= w;
        const int r = clampi(j - 1, 0, h - 1), c = clampi(i - 1, 0, w - 1);
        d = (inner ? D : R)[(size_t)comp * plane + (size_t)(r + c) * hs + r];
    }
    return pad_flow(fl, comp, t, h, w, j, i) + d;
}

__global__ void k_psi_smooth(PadFlow fl, const double *__restrict__ duv, const double *__restrict__ ring,
                             size_t fstride, int h, int w, int hs, size_t plane, int zero, double a,
                             double am1, int mix, double beta, double hx, double hy,
                             double *__restrict__ psi_s, size_t ps_stride)
{
    const int ip = blockIdx.x * blockDim.x + threadIdx.x, jp = blockIdx.y, t = blockIdx.z;
    const int m = h + 2, n = w + 2;
    if (ip >= n) return;
    const double *D = duv + (size_t)t * fstride, *R = ring + (size_t)t * fstride;
    // one-sided on the padded edge, central inside
    const int il = ip == 0 ? 0 : ip - 1, ir = ip == n - 1 ? n - 1 : ip + 1;
    const int ju = jp == 0 ? 0 : jp - 1, jd = jp == m - 1 ? m - 1 : jp + 1;
    const bool cx = ip > 0 && ip < n - 1, cy = jp > 0 && jp < m - 1;
    double g[4];  // ux, uy, vx, vy
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        const double dx = psi_full(fl, D, R, comp, t, h, w, hs, plane, zero, jp, ir) -
                          psi_full(fl, D, R, comp, t, h, w, hs, plane, zero, jp, il);
        const double dy = psi_full(fl, D, R, comp, t, h, w, hs, plane, zero, jd, ip) -
                          psi_full(fl, D, R, comp, t, h, w, hs, plane, zero, ju, ip);
        g[2 * comp] = cx ? dx / (2 * hx) : dx / hx;
        g[2 * comp + 1] = cy ? dy / (2 * hy) : dy / hy;
    }
    double s = g[0] * g[0] + g[1] * g[1] + g[2] * g[2] + g[3] * g[3];
    if (s < 0) s = 0;
    double ps = a * frh_det_pow(s + 1e-5, am1);
    if (mix) ps = (1.0 - beta) + beta * ps;
    psi_s[(size_t)t * ps_stride + (size_t)jp * n + ip] = ps;
}

hipError_t launch_psi_smooth(const float *u32, const float *v32, size_t fs32, const double *u64,
                             const double *v64, const double *duv, const double *ring, size_t fstride,
                             int T, const Skew &s, int zero, double a_smooth, double gnc_beta, double hx,
                             double hy, double *psi_s, size_t ps_stride, hipStream_t st)
{
    PadFlow fl{u32, v32, fs32, u64, v64};
    const int mix = gnc_beta >= 0 && a_smooth > 0.0 && a_smooth < 1.0 && gnc_beta < 1.0;
    const dim3 g((s.w + 2 + 63) / 64, s.h + 2, T), b(64);
    hipLaunchKernelGGL(k_psi_smooth, g, b, 0, st, fl, duv, ring, fstride, s.h, s.w, s.hs, s.plane, zero,
                       a_smooth, a_smooth - 1.0, mix, gnc_beta, hx, hy, psi_s, ps_stride);
    return hipGetLastError();
}

// Lag-block coefficients with the psi_s weights (the operations of
// assemble_pixel in the oracle): per cell the four edge weights, cu, cv
// from the padded initial flow, then the data terms as in k_assemble; nine
// planes out.
template <int C, typename WT>
__global__ void k_assemble_nl(const double *__restrict__ J, const WT *__restrict__ Wsk, PadFlow fl,
                              const double *__restrict__ psi_s, size_t ps_stride,
                              const double *__restrict__ duv, double *__restrict__ coef, size_t fstride,
                              int h, int w, int hs, size_t plane, double ax, double ay, AData ad,
                              int psi_one, int zero_init)
{
    const int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x, t = blockIdx.y;
    const int c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    const size_t q = (size_t)d * hs + r;
    const int n = w + 2, jp = r + 1, ip = c + 1;
    const double *ps = psi_s + (size_t)t * ps_stride;
    const double pc = ps[(size_t)jp * n + ip];
    const int nj[4] = {jp, jp, jp + 1, jp - 1}, ni[4] = {ip - 1, ip + 1, ip, ip};  // left, right, down, up
    const double wd[4] = {ax, ax, ay, ay};
    double wn[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) wn[k] = 0.5 * (pc + ps[(size_t)nj[k] * n + ni[k]]) * wd[k];
    const double uc = pad_flow(fl, 0, t, h, w, jp, ip), vc = pad_flow(fl, 1, t, h, w, jp, ip);
    double cu = wn[0] * (pad_flow(fl, 0, t, h, w, nj[0], ni[0]) - uc);
    double cv = wn[0] * (pad_flow(fl, 1, t, h, w, nj[0], ni[0]) - vc);
#pragma unroll
    for (int k = 1; k < 4; ++k) {
        cu = cu + wn[k] * (pad_flow(fl, 0, t, h, w, nj[k], ni[k]) - uc);
        cv = cv + wn[k] * (pad_flow(fl, 1, t, h, w, nj[k], ni[k]) - vc);
    }
    const double *Jt = J + (size_t)t * fstride;
    double x = 0.0, y = 0.0;
    if (!psi_one && !zero_init) {
        x = duv[(size_t)t * fstride + q];
        y = duv[(size_t)t * fstride + plane + q];
    }
    double Jr[6 * C], Wr[C], K[5];
#pragma unroll
    for (int e = 0; e < 6 * C; ++e) Jr[e] = Jt[(size_t)e * plane + q];
#pragma unroll
    for (int k = 0; k < C; ++k) Wr[k] = (double)Wsk[(size_t)k * plane + q];
    coef_from_raw_ws<C>(Jr, Wr, cu, cv, x, y, ((wn[0] + wn[1]) + wn[2]) + wn[3], ad, psi_one, K);
    double *co = coef + (size_t)t * fstride;
#pragma unroll
    for (int e = 0; e < 5; ++e) co[e * plane + q] = K[e];
#pragma unroll
    for (int k = 0; k < 4; ++k) co[(5 + k) * plane + q] = wn[k];
}

hipError_t launch_assemble_nl(const double *J, const void *Wsk, int w_f32, const float *u32, const float *v32,
                              size_t fs32, const double *u64, const double *v64, const double *psi_s,
                              size_t ps_stride, const double *duv, double *coef, size_t frame_stride, int T,
                              int C, const Skew &s, double ax, double ay, const double *a_data_minus1,
                              const double *a_data, int psi_one, int zero_init, double gnc_beta,
                              hipStream_t st)
{
    const AData ad = make_adata(C, a_data, a_data_minus1, gnc_beta);
    PadFlow fl{u32, v32, fs32, u64, v64};
    dim3 g(s.h + s.w - 1, T, (s.hs + 1023) / 1024), b(s.hs < 1024 ? s.hs : 1024);
#define FRH_ASMNL(CC)                                                                               \
    if (w_f32)                                                                                      \
        hipLaunchKernelGGL((k_assemble_nl<CC, float>), g, b, 0, st, J, (const float *)Wsk, fl, psi_s, ps_stride, \
                           duv, coef, frame_stride, s.h, s.w, s.hs, s.plane, ax, ay, ad, psi_one, zero_init); \
    else                                                                                            \
        hipLaunchKernelGGL((k_assemble_nl<CC, double>), g, b, 0, st, J, (const double *)Wsk, fl, psi_s,     \
                           ps_stride, duv, coef, frame_stride, s.h, s.w, s.hs, s.plane, ax, ay, ad, psi_one, \
                           zero_init)
    switch (C) {
    case 1: FRH_ASMNL(1); break;
    case 2: FRH_ASMNL(2); break;
    case 3: FRH_ASMNL(3); break;
    case 4: FRH_ASMNL(4); break;
    case 5: FRH_ASMNL(5); break;
    case 6: FRH_ASMNL(6); break;
    case 7: FRH_ASMNL(7); break;
    case 8: FRH_ASMNL(8); break;
    default: return hipErrorInvalidValue;
    }
#undef FRH_ASMNL
    return hipGetLastError();
}

// ------------------------------------------------------------- SOR sweeps
// core/level_solver.py:257-364: P consecutive sweeps of the column-major
// lexicographic SOR for one frame per work-group, lane r = interior row r.
// Cell (r, c) of sweep p is updated at wavefront step t = r + c + 2p: its
// up and left neighbours already hold sweep-p values and its down and right
// neighbours still hold sweep-(p-1) values, exactly as in the sequential
// column-by-column loop, so the result is identical to that loop.
//
// Registers per lane: the coefficients of the last 2P-1 cells the lane
// entered (a ring indexed at compile time: the step loop is unrolled over
// one ring period), the last two results of every sweep, and the pass-start
// values of the next cells (prefetched two steps ahead).  Row neighbours
// move between lanes with DPP wave shifts (out-of-wave lanes read 0) and
// between waves through LDS, with one barrier per step that waits for LDS
// traffic only.  Values of cells outside the level are 0; the weight of an
// out-of-level neighbour (which holds the cell's own start value in the
// reference) is folded into the cell's own term, see sor_pixel_coef in the
// oracle, which performs the same operations in the same order.
struct SweepArgs {
    const double *coef;   // 5 coefficient planes (unfused passes)
    double *duv;
    double *ring;
    size_t fstride, plane;
    int h, w, hs;         // h: rows of this launch (a band of the level in band mode)
    double ax, ay;
    int zero_init;
    int r0, hlev;         // band mode: the band's first row and the level's rows
    // fused passes: coefficients formed on entry from the raw planes
    const double *J, *Wsk, *cuv;
    AData ad;
    int psi_one;
    // chain of work-groups (k_sweep<1, 512, 0, false, true>): one per sweep
    // and frame; progress flags, failure word, sweeps, frames, padded
    // frames, publication interval
    int *cflags, *cerr;
    int cn, cT, cTpad, cpub;
    int cnb;    // band mode: bands per sweep (stages = cn x cnb, sweep-major)
    int cskip;  // the last stage stores its first sweep's results
};

// A work-group's place in a chain: its predecessor's and its own progress
// flag (null: none), the predecessor's progress it has acquired, the flag as
// loaded one step earlier, steps to the next publication, diagonals, and
// whether its start values are zero (the level's first pass, stage 0).
struct ChainState {
    int *fl_in, *fl_out;
    int avail, fpre, pcount, nd, zero_init;
    // band mode: the band above in the same sweep (its new bottom row) and
    // the band below in the previous sweep (its top row's start values)
    int *fl_top = nullptr, *fl_bot = nullptr;
    int avail_top = 0, fpre_top = 0, nd_top = 0, avail_bot = 0, fpre_bot = 0, nd_bot = 0;
    // a two-sweep stage that stores its first sweep's results (the last
    // stage of a pass with an odd number of sweeps)
    int skip = 0;
};

// C = 0: the pass reads precomputed coefficients; C >= 1: the pass forms
// each cell's coefficients (coef_from_raw, the same operations as
// k_assemble) when the cell enters the wavefront, from its J / W / cu, cv
// planes and start values.
//
// Register rings.  The coefficients of diagonal x live in K[x % NK] and the
// pass-start values of diagonal x in Lu/Lv[x % NL].  With C = 0 the
// coefficients of diagonal t + DK and the start values of diagonal t + DL
// are loaded at step t straight into their ring slots, so a load has DK
// (DL - 2 for the value handed across waves) steps to arrive before it is
// first used and no register copy forces an earlier wait.  The step loop is
// unrolled over U = lcm(NK, NL) steps so every slot index is static.
constexpr int sweep_gcd(int a, int b) { return b == 0 ? a : sweep_gcd(b, a % b); }
constexpr int sweep_lcm(int a, int b) { return a / sweep_gcd(a, b) * b; }
constexpr int kSweepNL = 6, kSweepDL = 5;
// smallest ring size >= nmin whose unroll period lcm(n, nl) is at most lmax
constexpr int sweep_ring(int nmin, int nl = kSweepNL, int lmax = 18)
{
    int n = nmin;
    while (sweep_lcm(n, nl) > lmax) ++n;
    return n;
}
// 1024-lane work-groups (4 waves per SIMD, 128 registers per lane): shorter
// prefetch distances (coefficients 2 diagonals ahead, start values 3), which
// keep the 2-sweep form within 128 registers (8 B per lane spilled instead of
// 76 with the 6-slot rings; 24% less time per pass)
constexpr int kSweepNLTight = 4, kSweepDLTight = 3;

// C = kSmoothNL: nine precomputed coefficients per cell (per-cell smoothness
// weights, a_smooth != 1).
constexpr int kSmoothNL = -1;

// Band mode (levels of more than 1024 rows, one sweep per pass): a level is
// swept in bands of at most 512 rows, one launch per band from the top.
// One sweep's new values depend only on the new values of the cell above and
// of the cell to the left, so finishing a band before starting the next
// gives the column-major loop's values: the band's top row takes the new
// values of the row above from the previous band's launch, and its bottom
// row the start values of the row below, which the next band has not
// touched yet.  Both come from memory through a small ring, kBandNR slots,
// loaded kBandDR steps ahead.
constexpr int kBandNR = 4, kBandDR = 3;

template <int P, int C, bool TIGHT = false>
struct SweepState {
    static constexpr int NR = 2 * P - 1;  // diagonals a lane's P sweeps span
    static constexpr int NL = TIGHT ? kSweepNLTight : kSweepNL, DL = TIGHT ? kSweepDLTight : kSweepDL;
    static constexpr int NK = TIGHT ? sweep_ring(NR + 2, NL, 20) : sweep_ring(C <= 0 ? NR + 3 : NR);
    static constexpr int DK = NK - NR;    // coefficient prefetch distance (C <= 0)
    static constexpr int U = sweep_lcm(NK, NL);
    static constexpr int NRAW = C > 0 ? 7 * C + 2 : 1;
    static constexpr int NKC = C == kSmoothNL ? 9 : 5;  // coefficients per cell
    double K[NK][NKC];
    double Ru[2][P];      // results of step t-1 / t-2, slot by step parity
    double Rv[2][P];
    double Lu[NL], Lv[NL];
    double raw[NRAW];     // raw data of diag t+1 loaded at step t, C >= 1
    double Bu[kBandNR], Bv[kBandNR];  // band mode: neighbour rows of the band's edge lanes
};

// Band mode: byte offset of the value the band's top lane (new value of the
// row above) or bottom lane (start value of the row below) needs at step t,
// or kOutOfRange.
__device__ inline unsigned band_offset(const SweepArgs &a, int r, int t, int zero_init)
{
    const int h = a.h, w = a.w;
    unsigned q = kOutOfRange;
    if (r == 0 && a.r0 > 0) {
        const int c = t;
        if ((unsigned)c < (unsigned)w) q = ((unsigned)(a.r0 - 1 + c) * (unsigned)a.hs + (unsigned)(a.r0 - 1)) * 8u;
    } else if (r == h - 1 && a.r0 + h < a.hlev && !zero_init) {
        // (on a level's first pass every start value is 0: the memory still
        // holds the previous level's values)
        const int c = t - (h - 1);
        if ((unsigned)c < (unsigned)w) q = ((unsigned)(a.r0 + h + c) * (unsigned)a.hs + (unsigned)(a.r0 + h)) * 8u;
    }
    return q;
}

template <int C>
__device__ __forceinline__ void load_raw(double *raw, const SweepArgs &a, size_t fo, size_t q)
{
    const double *J = a.J + fo, *cuv = a.cuv + fo;
#pragma unroll
    for (int e = 0; e < 6 * C; ++e) raw[e] = J[(size_t)e * a.plane + q];
#pragma unroll
    for (int k = 0; k < C; ++k) raw[6 * C + k] = a.Wsk[(size_t)k * a.plane + q];
    raw[7 * C] = cuv[q];
    raw[7 * C + 1] = cuv[a.plane + q];
}

struct SweepBufs {
    __amdgpu_buffer_rsrc_t coef, duv;  // the 5 coefficient planes; du, dv
    unsigned plane_bytes;
};

// Chain waits before step t's loads (t = -1: the prologue's): the start
// values of diagonal t + DL from the predecessor sweep of the same rows,
// which stores diagonal x at its step x; in band mode (bands of kBandRows
// rows) also the new values of the row above (column t + kBandDR, the band
// above's diagonal t + kBandDR + kBandRows - 1) and the start values of the
// row below (column t + kBandDR - (h - 1), the lower band's diagonal of the
// same number, from the previous sweep).
constexpr int kBandRows = 512;

template <int DL>
__device__ __forceinline__ void chain_wait(ChainState &cs, const SweepArgs &a, int t)
{
    if (cs.fl_in) {
        const int need = min(t + DL + 1, cs.nd);
        if (cs.avail < need) cs.avail = chain_acquire(cs.fl_in, need, __builtin_amdgcn_readfirstlane(cs.fpre), a.cerr);
        cs.fpre = __hip_atomic_load(cs.fl_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (cs.fl_top) {
        const int need = min(t + kBandDR + kBandRows, cs.nd_top);
        if (cs.avail_top < need)
            cs.avail_top = chain_acquire(cs.fl_top, need, __builtin_amdgcn_readfirstlane(cs.fpre_top), a.cerr);
        cs.fpre_top = __hip_atomic_load(cs.fl_top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (cs.fl_bot) {
        const int need = min(t + kBandDR - a.h + 2, cs.nd_bot);
        if (cs.avail_bot < need)
            cs.avail_bot = chain_acquire(cs.fl_bot, need, __builtin_amdgcn_readfirstlane(cs.fpre_bot), a.cerr);
        cs.fpre_bot = __hip_atomic_load(cs.fl_bot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int P, int MAXW, int C, int S, bool TT, bool BAND, bool CH>
__device__ __forceinline__ void sweep_step(SweepState<P, C, TT> &st, const SweepArgs &a, SweepBufs &bf,
                                           double *rdu, double *rdv,
                                           double (*xb)[MAXW][2][P + 1][2], int t, int r,
                                           int lane, int wv, int nw, int Tend, size_t fo, ChainState &cs)
{
    using St = SweepState<P, C, TT>;
    constexpr int NK = St::NK, NL = St::NL, DK = St::DK, DL = St::DL;
    static_assert(!BAND || (P == 1 && C <= 0 && St::U % kBandNR == 0), "band mode: one sweep per pass");
    // the lane's row and the diagonal offset in the level (band mode: the
    // launch's row r is level row r + R0, its diagonal d is level diagonal d + R0)
    const int R0 = BAND ? a.r0 : 0;
    // Every step, also the ones past the last diagonal (t > Tend, where no
    // lane has a cell), issues the same vector memory instructions, with no
    // branch around them, so the compiler can count the loads in flight and
    // wait for each one only where its value is first used.  Lanes without a
    // cell use an out-of-range buffer offset: their loads return 0 (a cell
    // outside the level has zero coefficients and value) and their stores
    // are dropped, without touching memory.
    (void)Tend;
    const int h = a.h, w = a.w, hs = a.hs;
    const bool row_ok = r < h;
    const int zero_init = CH ? cs.zero_init : a.zero_init;
    if constexpr (CH) chain_wait<DL>(cs, a, t);
    {
        if constexpr (C <= 0) {
            const int d = t + DK;
            const unsigned q = slot_offset(row_ok, d + R0, r + R0, w, hs);
            double *k = st.K[(S + DK) % NK];
#pragma unroll
            for (int e = 0; e < St::NKC; ++e) k[e] = buf_load_f64(bf.coef, q, e * bf.plane_bytes);
        } else {
            const int d = t + 1, c = d - r;
            if (row_ok && c >= 0 && c < w) {
                load_raw<C>(st.raw, a, fo, (size_t)d * hs + r);
            } else {
#pragma unroll
                for (int e = 0; e < 7 * C + 2; ++e) st.raw[e] = 0.0;
            }
        }
        const int d = t + DL;
        const unsigned q = zero_init ? kOutOfRange : slot_offset(row_ok, d + R0, r + R0, w, hs);
        st.Lu[(S + DL) % NL] = buf_load_f64(bf.duv, q, 0);
        st.Lv[(S + DL) % NL] = buf_load_f64(bf.duv, q, bf.plane_bytes);
        if constexpr (BAND) {
            const unsigned qb = band_offset(a, r, t + kBandDR, zero_init);
            st.Bu[(S + kBandDR) % kBandNR] = buf_load_f64(bf.duv, qb, 0);
            st.Bv[(S + kBandDR) % kBandNR] = buf_load_f64(bf.duv, qb, bf.plane_bytes);
        }
    }
    const double lpu = st.Lu[S % NL], lpv = st.Lv[S % NL];              // own start value
    const double ldu = st.Lu[(S + 1) % NL], ldv = st.Lv[(S + 1) % NL];  // right neighbour's
    constexpr int PV = (S + 1) & 1;  // results of step t-1
    constexpr int P2 = S & 1;        // results of step t-2 (overwritten below)
    double upu[P], upv[P], dnu[P], dnv[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        upu[p] = wave_from_prev_lane(st.Ru[PV][p]);
        upv[p] = wave_from_prev_lane(st.Rv[PV][p]);
        dnu[p] = wave_from_next_lane(p == 0 ? ldu : st.Ru[PV][p > 0 ? p - 1 : 0]);
        dnv[p] = wave_from_next_lane(p == 0 ? ldv : st.Rv[PV][p > 0 ? p - 1 : 0]);
    }
    const int pb = (t + 1) & 1;
    if (lane == 0 && wv > 0) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            upu[p] = xb[pb][wv - 1][1][p][0];
            upv[p] = xb[pb][wv - 1][1][p][1];
        }
    }
    if (lane == 63 && wv < nw - 1) {
        dnu[0] = xb[pb][wv + 1][0][P][0];
        dnv[0] = xb[pb][wv + 1][0][P][1];
#pragma unroll
        for (int p = 1; p < P; ++p) {
            dnu[p] = xb[pb][wv + 1][0][p - 1][0];
            dnv[p] = xb[pb][wv + 1][0][p - 1][1];
        }
    }
    if constexpr (BAND) {
        // the rows above and below the band (0 at the level's own edges)
        if ((r == 0 && a.r0 > 0) || (r == h - 1 && a.r0 + h < a.hlev)) {
            upu[0] = r == 0 ? st.Bu[S % kBandNR] : upu[0];
            upv[0] = r == 0 ? st.Bv[S % kBandNR] : upv[0];
            dnu[0] = r == 0 ? dnu[0] : st.Bu[S % kBandNR];
            dnv[0] = r == 0 ? dnv[0] : st.Bv[S % kBandNR];
        }
    }
    const double ax = a.ax, ay = a.ay, om1 = 1 - 1.95;
    // out-of-level neighbour weight b = ax * (#left/right outside) + ay * (#up/down outside)
    const int rg = r + R0, HL = BAND ? a.hlev : h;  // level row, level rows
    const double brow = ay * (double)((rg == 0) + (rg == HL - 1));
    const int rlo = wv * 64, rhi = (rlo + 63 < h - 1 ? rlo + 63 : h - 1);
    double cu[P], cv[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        cu[p] = 0.0;
        cv[p] = 0.0;
        // uniform: does any lane of this wave own a cell of sweep p?
        if (t - 2 * p - rlo >= 0 && t - 2 * p - rhi <= w - 1) {
            const int c = t - r - 2 * p;
            const double *Kp = st.K[((S - 2 * p) % NK + NK) % NK];
            double ou = p == 0 ? lpu : st.Ru[P2][p > 0 ? p - 1 : 0];
            double ov = p == 0 ? lpv : st.Rv[P2][p > 0 ? p - 1 : 0];
            double ru = p == 0 ? ldu : st.Ru[PV][p > 0 ? p - 1 : 0];
            double rv = p == 0 ? ldv : st.Rv[PV][p > 0 ? p - 1 : 0];
            double lu = st.Ru[PV][p], lv = st.Rv[PV][p];
            double A12 = Kp[0], Bu = Kp[1], Bv = Kp[2], wu = Kp[3], wvv = Kp[4];
            double b, nu, nv;
            if constexpr (C == kSmoothNL) {
                // per-cell edge weights; an out-of-level neighbour's weight
                // moves to the cell's own term, summed left, right, up, down
                const double wl = Kp[5], wr = Kp[6], wdn = Kp[7], wup = Kp[8];
                b = (((c == 0 ? wl : 0.0) + (c == w - 1 ? wr : 0.0)) + (rg == 0 ? wup : 0.0)) +
                    (rg == HL - 1 ? wdn : 0.0);
                nu = fma(wup, upu[p], fma(wdn, dnu[p], fma(wr, ru, fma(wl, lu, Bu))));
                nv = fma(wup, upv[p], fma(wdn, dnv[p], fma(wr, rv, fma(wl, lv, Bv))));
            } else {
                b = fma(ax, (double)((c == 0) + (c == w - 1)), brow);
                nu = fma(ax, lu + ru, Bu);
                nu = fma(ay, dnu[p] + upu[p], nu);
                nv = fma(ax, lv + rv, Bv);
                nv = fma(ay, dnv[p] + upv[p], nv);
            }
            nu = fma(-A12, ov, nu);
            double dun = fma(wu, nu, fma(wu, b, om1) * ou);
            nv = fma(-A12, dun, nv);
            double dvn = fma(wvv, nv, fma(wvv, b, om1) * ov);
            const bool ok = row_ok && (unsigned)c < (unsigned)w;
            if constexpr (C <= 0) {
                // a cell outside the level has zero coefficients and a zero
                // start value, so it computes 0 (up to sign)
                cu[p] = dun;
                cv[p] = dvn;
            } else {
                cu[p] = ok ? dun : 0.0;
                cv[p] = ok ? dvn : 0.0;
            }
            if (p == P - 1 && rdu && ok && (rg == 0 || rg == HL - 1 || c == 0 || c == w - 1)) {
                size_t q = (size_t)(t - 2 * (P - 1) + R0) * hs + rg;
                rdu[q] = ou;
                rdv[q] = ov;
            }
        }
    }
    {
        // sweep P-1 leaves diag t - 2(P-1): store its results
        const unsigned q = slot_offset(row_ok, t - 2 * (P - 1) + R0, rg, w, hs);
        double su = cu[P - 1], sv = cv[P - 1];
        if constexpr (CH && P >= 2) {
            // a stage that skips its last sweep: sweep P - 2's results of the
            // same diagonal, from step t - 2
            su = cs.skip ? st.Ru[P2][P - 2] : su;
            sv = cs.skip ? st.Rv[P2][P - 2] : sv;
        }
        // (the pass's results are read again only by the next pass or the
        // assembly, after every frame's pass: non-temporal stores)
        buf_store_f64<kNT>(su, bf.duv, q, 0);
        buf_store_f64<kNT>(sv, bf.duv, q, bf.plane_bytes);
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        st.Ru[P2][p] = cu[p];
        st.Rv[P2][p] = cv[p];
    }
    if constexpr (C > 0) {
        // coefficients of the cell entering at step t+1 (diag t+1), from its
        // raw data and its start values
        coef_from_raw<C>(st.raw, st.raw + 6 * C, st.raw[7 * C], st.raw[7 * C + 1], ldu, ldv, ax, ay,
                         a.ad, a.psi_one, st.K[(S + 1) % NK]);
    }
    const int wb = t & 1;
    if (lane == 0) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xb[wb][wv][0][p][0] = cu[p];
            xb[wb][wv][0][p][1] = cv[p];
        }
        // start value of diag t+2: the upper neighbour wave's p = 0 down value at step t+1
        xb[wb][wv][0][P][0] = st.Lu[(S + 2) % NL];
        xb[wb][wv][0][P][1] = st.Lv[(S + 2) % NL];
    }
    if (lane == 63) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xb[wb][wv][1][p][0] = cu[p];
            xb[wb][wv][1][p][1] = cv[p];
        }
    }
    if constexpr (CH) {
        // publish the number of diagonals stored (those before t - 2(P - 1)
        // + 1) every cpub steps and after the last diagonal, once this
        // step's stores of every wave have been performed
        const int stored = t - 2 * (P - 1) + 1;
        const bool pub = cs.fl_out && (--cs.pcount == 0 || stored == cs.nd);
        if (cs.pcount == 0) cs.pcount = a.cpub;
        if (pub) wait_stores();
        lds_barrier();
        if (pub && r == 0)
            __hip_atomic_store(cs.fl_out, max(min(stored, cs.nd), 0), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        lds_barrier();
    }
    __builtin_amdgcn_sched_barrier(0);  // keep each step's loads within the step
}

template <int P, int MAXW, int C, bool TT, bool BAND, bool CH, int... S>
__device__ __forceinline__ void sweep_block(SweepState<P, C, TT> &st, const SweepArgs &a,
                                            SweepBufs &bf, double *rdu,
                                            double *rdv, double (*xb)[MAXW][2][P + 1][2], int t0,
                                            int r, int lane, int wv, int nw, int Tend, size_t fo, ChainState &cs,
                                            std::integer_sequence<int, S...>)
{
    (sweep_step<P, MAXW, C, S, TT, BAND, CH>(st, a, bf, rdu, rdv, xb, t0 + S, r, lane, wv, nw, Tend, fo, cs),
     ...);
}

// One wavefront pass over frame f by the work-group's first nthr threads
// (a multiple of 64; threads past them have left the kernel).
template <int P, int MAXT, int C, bool BAND, bool CH>
__device__ __forceinline__ void sweep_frame_body(const SweepArgs &a, int f, int nthr,
                                                 double (*xb)[MAXT / 64][2][P + 1][2], ChainState &cs, bool record)
{
    constexpr bool TT = MAXT > 512 && C == 0 && P >= 2;  // the 1-sweep form fits without it
    using St = SweepState<P, C, TT>;
    constexpr int NK = St::NK, NL = St::NL, DK = St::DK, DL = St::DL, U = St::U;
    constexpr int MAXW = MAXT / 64;
    const int r = threadIdx.x;
    const int lane = r & 63;
    const int wv = __builtin_amdgcn_readfirstlane(r >> 6);
    const int nw = nthr >> 6;
    const size_t fo = (size_t)f * a.fstride;
    const double *cb = a.coef + fo;
    double *du = a.duv + fo;
    double *rdu = record ? a.ring + fo : nullptr;
    double *rdv = record ? rdu + a.plane : nullptr;
    for (int k = threadIdx.x; k < 2 * MAXW * 2 * (P + 1) * 2; k += nthr)
        (&xb[0][0][0][0][0])[k] = 0.0;
    const int h = a.h, w = a.w, hs = a.hs;
    const bool row_ok = r < h;
    St st;
#pragma unroll
    for (int k = 0; k < NK; ++k)
#pragma unroll
        for (int e = 0; e < St::NKC; ++e) st.K[k][e] = 0.0;
#pragma unroll
    for (int p = 0; p < P; ++p) st.Ru[0][p] = st.Ru[1][p] = st.Rv[0][p] = st.Rv[1][p] = 0.0;
#pragma unroll
    for (int k = 0; k < NL; ++k) st.Lu[k] = st.Lv[k] = 0.0;
#pragma unroll
    for (int e = 0; e < St::NRAW; ++e) st.raw[e] = 0.0;
    // prologue: start values of diags 0 .. DL-1; coefficients of diags
    // 0 .. DK-1 (C = 0) or of diag 0 (C >= 1)
    auto cell_ok = [&](int d) { int c = d - r; return row_ok && c >= 0 && c < w; };
    const int R0 = BAND ? a.r0 : 0;
    SweepBufs bf;
    bf.plane_bytes = (unsigned)(a.plane * sizeof(double));
    bf.coef = make_rsrc(cb, C <= 0 ? St::NKC * bf.plane_bytes : 0u);
    bf.duv = make_rsrc(du, 2 * bf.plane_bytes);
    if constexpr (CH) chain_wait<DL>(cs, a, -1);
#pragma unroll
    for (int d = 0; d < DL; ++d) {
        const unsigned q = cs.zero_init ? kOutOfRange : slot_offset(row_ok, d + R0, r + R0, w, hs);
        st.Lu[d] = buf_load_f64(bf.duv, q, 0);
        st.Lv[d] = buf_load_f64(bf.duv, q, bf.plane_bytes);
    }
    if constexpr (BAND) {
#pragma unroll
        for (int k = 0; k < kBandDR; ++k) {
            const unsigned qb = band_offset(a, r, k, cs.zero_init);
            st.Bu[k] = buf_load_f64(bf.duv, qb, 0);
            st.Bv[k] = buf_load_f64(bf.duv, qb, bf.plane_bytes);
        }
    }
    if constexpr (C <= 0) {
#pragma unroll
        for (int d = 0; d < DK; ++d) {
            const unsigned q = slot_offset(row_ok, d + R0, r + R0, w, hs);
#pragma unroll
            for (int e = 0; e < St::NKC; ++e) st.K[d][e] = buf_load_f64(bf.coef, q, e * bf.plane_bytes);
        }
    } else {
        if (cell_ok(0)) {
            double raw0[7 * C + 2];
            load_raw<C>(raw0, a, fo, (size_t)r);
            coef_from_raw<C>(raw0, raw0 + 6 * C, raw0[7 * C], raw0[7 * C + 1], st.Lu[0], st.Lv[0],
                             a.ax, a.ay, a.ad, a.psi_one, st.K[0]);
        }
    }
    __syncthreads();
    const int Tend = (h + w - 2) + 2 * (P - 1);
    for (int t0 = 0; t0 <= Tend; t0 += U)
        sweep_block<P, MAXW, C, TT, BAND, CH>(st, a, bf, rdu, rdv, xb, t0, r, lane, wv, nw, Tend, fo, cs,
                                              std::make_integer_sequence<int, U>{});
}

template <int P, int MAXT, int C, bool BAND, bool CH = false>
__device__ __forceinline__ void sweep_frame(const SweepArgs &a, int f, int nthr,
                                            double (*xb)[MAXT / 64][2][P + 1][2])
{
    static_assert(!CH || (BAND ? (P == 1 && C <= 0) : (P <= 3 && C == 0)),
                  "a chain stage runs one sweep (up to three without bands)");
    ChainState cs{nullptr, nullptr, 0, 0, 0, 0, a.zero_init};
    bool record = a.ring != nullptr;  // border ring: the pass's last sweep
    if constexpr (CH) {
        // work-groups numbered stage-major: stage = sweep j of the pass (and
        // band b of the level in band mode: stage j nb + b)
        const int stage = f / a.cTpad, nb = BAND ? a.cnb : 1;
        f -= stage * a.cTpad;
        if (f >= a.cT) return;
        const int j = stage / nb, b = stage - j * nb;
        auto flag = [&](int sj, int sb) {
            return a.cflags + (size_t)((sj * nb + sb) * a.cTpad + f) * kChainFlagStride;
        };
        cs.pcount = a.cpub;
        cs.zero_init = a.zero_init && j == 0;
        cs.skip = a.cskip && j == a.cn - 1;
        if (j > 0) cs.fl_in = flag(j - 1, b);
        if (j < a.cn - 1 || b < nb - 1) cs.fl_out = flag(j, b);
        record = record && j == a.cn - 1;
        if constexpr (BAND) {
            SweepArgs ab = a;
            ab.r0 = b * kBandRows;
            ab.h = a.hlev - ab.r0 < kBandRows ? a.hlev - ab.r0 : kBandRows;
            cs.nd = ab.h + a.w - 1;
            if (b > 0) {
                cs.fl_top = flag(j, b - 1);
                cs.nd_top = kBandRows + a.w - 1;
            }
            if (j > 0 && b < nb - 1) {
                cs.fl_bot = flag(j - 1, b + 1);
                const int hb = a.hlev - ab.r0 - kBandRows;
                cs.nd_bot = (hb < kBandRows ? hb : kBandRows) + a.w - 1;
            }
            sweep_frame_body<P, MAXT, C, BAND, CH>(ab, f, nthr, xb, cs, record);
            return;
        }
        cs.nd = a.h + a.w - 1;
    }
    sweep_frame_body<P, MAXT, C, BAND, CH>(a, f, nthr, xb, cs, record);
}

template <int P, int MAXT, int C, bool BAND = false, bool CH = false>
__global__ __launch_bounds__(MAXT) void k_sweep(SweepArgs a)
{
    __shared__ double xb[2][MAXT / 64][2][P + 1][2];
    sweep_frame<P, MAXT, C, BAND, CH>(a, blockIdx.x, blockDim.x, xb);
}

template <int P, int C>
static hipError_t launch_sweep_pc(SweepArgs a, int T, hipStream_t st)
{
    // buffer offsets are 32-bit and must stay below kOutOfRange
    if (a.plane * sizeof(double) * SweepState<P, C>::NKC >= kOutOfRange) return hipErrorInvalidValue;
    a.r0 = 0;
    a.hlev = a.h;
    if (a.hs <= 512) {
        hipLaunchKernelGGL((k_sweep<P, 512, C>), dim3(T), dim3(a.hs), 0, st, a);
    } else if (a.hs <= 1024 && P <= 2) {
        hipLaunchKernelGGL((k_sweep<(P <= 2 ? P : 2), 1024, C>), dim3(T), dim3(a.hs), 0, st, a);
    } else if (a.hs > 1024 && P == 1 && C <= 0) {
        // band mode: bands of 512 rows from the top, one launch each (512-lane
        // work-groups keep the band rings in registers)
        const int hl = a.h;
        for (int r0 = 0; r0 < hl; r0 += 512) {
            SweepArgs b = a;
            b.r0 = r0;
            b.h = hl - r0 < 512 ? hl - r0 : 512;
            hipLaunchKernelGGL((k_sweep<1, 512, (C <= 0 ? C : 0), true>), dim3(T), dim3(64 * ((b.h + 63) / 64)), 0,
                               st, b);
            const hipError_t e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <int P>
static hipError_t launch_sweep_p(const SweepArgs &a, int C, int T, hipStream_t st)
{
    switch (C) {
    case 0: return launch_sweep_pc<P, 0>(a, T, st);
    case 1: return launch_sweep_pc<P, 1>(a, T, st);
    case 2: return launch_sweep_pc<P, 2>(a, T, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_sweep(int P, const double *coef, double *duv, double *ring, size_t frame_stride,
                        int T, const Skew &s, double ax, double ay, int zero_init, hipStream_t st)
{
    SweepArgs a{};
    a.coef = coef; a.duv = duv; a.ring = ring; a.fstride = frame_stride; a.plane = s.plane;
    a.h = s.h; a.w = s.w; a.hs = s.hs; a.ax = ax; a.ay = ay; a.zero_init = zero_init;
    switch (P) {
    case 1: return launch_sweep_p<1>(a, 0, T, st);
    case 2: return launch_sweep_p<2>(a, 0, T, st);
    case 3: return launch_sweep_p<3>(a, 0, T, st);
    case 4: return launch_sweep_p<4>(a, 0, T, st);
    case 5: return launch_sweep_p<5>(a, 0, T, st);
    default: return hipErrorInvalidValue;
    }
}

// n sweeps of a pass as one chain of work-groups (levels of up to 512 rows,
// linear smoothness); flags: n * chain_tpad(T) progress flags.
hipError_t launch_sweep_chain(int n, int per_stage, const double *coef, double *duv, double *ring,
                              size_t frame_stride, int T, const Skew &s, double ax, double ay, int zero_init,
                              int *flags, int *err, hipStream_t st)
{
    SweepArgs a{};
    a.coef = coef; a.duv = duv; a.ring = ring; a.fstride = frame_stride; a.plane = s.plane;
    a.h = s.h; a.w = s.w; a.hs = s.hs; a.ax = ax; a.ay = ay; a.zero_init = zero_init;
    a.r0 = 0;
    a.hlev = a.h;
    // stages of per_stage (1 to 3) sweeps; when n leaves the last stage one
    // sweep short, that stage stores its second-to-last sweep's results
    if (per_stage < 1 || per_stage > 3 || n < 1 || (per_stage > 1 && n % per_stage != 0 &&
                                                    n % per_stage != per_stage - 1))
        return hipErrorInvalidValue;
    const int stages = (n + per_stage - 1) / per_stage;
    a.cflags = flags;
    a.cerr = err;
    a.cn = stages;
    a.cskip = per_stage > 1 && n % per_stage != 0;
    a.cT = T;
    a.cTpad = chain_tpad(T);
    a.cpub = chain_publish_interval(s.h + s.w - 1);
    if (s.hs > 512 ||
        a.plane * sizeof(double) * kCoefPlanes >= kOutOfRange || (size_t)stages * a.cTpad > 0x7fffffff)
        return hipErrorInvalidValue;
    hipError_t e = launch_clear_flags(flags, stages * a.cTpad, st);
    if (e != hipSuccess) return e;
    if (per_stage == 3)
        hipLaunchKernelGGL((k_sweep<3, 512, 0, false, true>), dim3(stages * a.cTpad), dim3(a.hs), 0, st, a);
    else if (per_stage == 2)
        hipLaunchKernelGGL((k_sweep<2, 512, 0, false, true>), dim3(stages * a.cTpad), dim3(a.hs), 0, st, a);
    else
        hipLaunchKernelGGL((k_sweep<1, 512, 0, false, true>), dim3(stages * a.cTpad), dim3(a.hs), 0, st, a);
    return hipGetLastError();
}

// n sweeps of a level of more than 1024 rows (band mode) as one chain of
// work-groups: stages (sweep j, band b of kBandRows rows), sweep-major, each
// waiting for the band above in its sweep and for its own and the lower
// band's rows in the previous sweep; nl: nine coefficients per cell
// (a_smooth != 1).  flags: n * bands * chain_tpad(T) progress flags.
hipError_t launch_sweep_band_chain(int n, int nl, const double *coef, double *duv, double *ring,
                                   size_t frame_stride, int T, const Skew &s, double ax, double ay, int zero_init,
                                   int *flags, int *err, hipStream_t st)
{
    SweepArgs a{};
    a.coef = coef; a.duv = duv; a.ring = ring; a.fstride = frame_stride; a.plane = s.plane;
    a.h = s.h; a.w = s.w; a.hs = s.hs; a.ax = ax; a.ay = ay; a.zero_init = zero_init;
    a.r0 = 0;
    a.hlev = s.h;
    a.cflags = flags;
    a.cerr = err;
    a.cn = n;
    a.cnb = (s.h + kBandRows - 1) / kBandRows;
    a.cT = T;
    a.cTpad = chain_tpad(T);
    a.cpub = chain_publish_interval(kBandRows + s.w - 1);
    const size_t stages = (size_t)n * a.cnb;
    if (n < 1 || s.hs <= 512 || a.plane * sizeof(double) * (nl ? 9 : kCoefPlanes) >= kOutOfRange ||
        stages * a.cTpad > 0x7fffffff)
        return hipErrorInvalidValue;
    hipError_t e = launch_clear_flags(flags, (int)(stages * a.cTpad), st);
    if (e != hipSuccess) return e;
    const dim3 g((unsigned)(stages * a.cTpad)), b(kBandRows);
    if (nl)
        hipLaunchKernelGGL((k_sweep<1, 512, kSmoothNL, true, true>), g, b, 0, st, a);
    else
        hipLaunchKernelGGL((k_sweep<1, 512, 0, true, true>), g, b, 0, st, a);
    return hipGetLastError();
}

// Work-groups of a sweep chain that run at once, one per compute unit (0:
// no device).  Two chain stages on one compute unit share its SIMDs and
// measured slower than the pipelined pass (64 frames: 96.9 against 53.8 ms
// of sweeps per step).
int sweep_chain_slots()
{
    int dev = 0, ncu = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void *)k_sweep<1, 512, 0, false, true>, 512, 0) !=
            hipSuccess ||
        per < 1)
        return 0;
    return ncu;
}

hipError_t launch_sweep_nl(int P, const double *coef, double *duv, double *ring, size_t frame_stride, int T,
                           const Skew &s, int zero_init, hipStream_t st)
{
    SweepArgs a{};
    a.coef = coef; a.duv = duv; a.ring = ring; a.fstride = frame_stride; a.plane = s.plane;
    a.h = s.h; a.w = s.w; a.hs = s.hs; a.zero_init = zero_init;
    switch (P) {
    case 1: return launch_sweep_pc<1, kSmoothNL>(a, T, st);
    case 2: return launch_sweep_pc<2, kSmoothNL>(a, T, st);
    default: return hipErrorInvalidValue;
    }
}

bool fused_sweep_supported(int C) { return C == 1 || C == 2; }

hipError_t launch_sweep_fused(int P, const double *J, const double *Wsk, const double *cuv,
                              double *duv, double *ring, size_t frame_stride, int T, int C,
                              const Skew &s, double ax, double ay, const double *a_data_minus1,
                              const double *a_data, int psi_one, int zero_init, double gnc_beta,
                              hipStream_t st)
{
    SweepArgs a{};
    a.coef = nullptr; a.duv = duv; a.ring = ring; a.fstride = frame_stride; a.plane = s.plane;
    a.h = s.h; a.w = s.w; a.hs = s.hs; a.ax = ax; a.ay = ay; a.zero_init = zero_init;
    a.J = J; a.Wsk = Wsk; a.cuv = cuv; a.psi_one = psi_one;
    a.ad = make_adata(C, a_data, a_data_minus1, gnc_beta);
    switch (P) {
    case 1: return launch_sweep_p<1>(a, C, T, st);
    case 2: return launch_sweep_p<2>(a, C, T, st);
    case 3: return launch_sweep_p<3>(a, C, T, st);
    case 4: return launch_sweep_p<4>(a, C, T, st);
    case 5: return launch_sweep_p<5>(a, C, T, st);
    default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------- median + u += du
// scipy.ndimage.median_filter(size=(5,5), mode="mirror") (core/optical_flow
// .py:607-609), then u = u + du (:610-611), with the selection network of
// frh_median.h.
// One wave per 64 consecutive rows of one diagonal of the skewed layout, so
// each of the 50 window loads is one contiguous run of slots.  Waves are
// numbered frame-major, diagonal, row chunk and dealt to the XCDs in
// contiguous ranges (neighbouring diagonals share the lines of the
// row-major flow planes).
// Rows staged per window diagonal by the median's interior path: the 64
// rows of a wave, 2 above, 2 below, and the second half of a 16-byte load
// past the end.
constexpr int kMedStageRows = 70;
using MedStage = double[2][9][kMedStageRows];

// Diagonal d, rows r0 .. r0 + 63 of one frame: du (and dv = du + plane) its
// skewed increments, ut / vt its row-major level flow; the sum goes to
// out_f32 / out_f64 (frame's (h, w, 2) interleaved output) or to the planes
// out_f32, out_f32 + h w.  stage: the wave's LDS (one-wave work-groups).
__device__ __forceinline__ void median_add_px(const double *__restrict__ du, int h, int w, int hs, size_t plane,
                                              int do_median, const float *__restrict__ ut,
                                              const float *__restrict__ vt, float *out_f32, double *out_f64,
                                              int out_interleaved, int d, int r0, MedStage &stage)
{
    const int r = r0 + threadIdx.x, c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    const double *dv = du + plane;
    double mu, mv;
    // rows of this wave that hold level cells, and their columns
    const int ra = max(r0, d - w + 1), rb = min(min(r0 + 63, d), h - 1);
    const bool interior = ra >= 2 && rb <= h - 3 && d - rb >= 2 && d - ra <= w - 3;
    if (do_median && interior) {
        // uniform: no window of the wave is mirrored.  Window sample (i, j)
        // of the cell in row r lies on diagonal d + k - 4 (k = i + j), row
        // r - 2 + i, so the wave's windows cover rows ra - 2 .. rb + 2 of
        // the nine diagonals d - 4 .. d + 4.  Each of these runs (per field)
        // is read once, in 16-byte pairs, into the wave's LDS, and every
        // lane takes its 25 samples per field from there.
        const __amdgpu_buffer_rsrc_t rs = make_rsrc(du, (unsigned)(2 * plane * sizeof(double)));
        const unsigned pb = (unsigned)(plane * sizeof(double));
        const int li = r - ra, m = rb - ra + 1, n = m + 4;  // lane's row in the run, active lanes, rows per run
        for (int q = li; 2 * q < n; q += m) {
#pragma unroll
            for (int k = 0; k <= 8; ++k) {
                const unsigned off = (unsigned)(((size_t)(d + k - 4) * hs + (ra - 2) + 2 * q) * sizeof(double));
                const F64x2 pu = buf_load_f64x2(rs, off, 0), pv = buf_load_f64x2(rs, off, pb);
                stage[0][k][2 * q] = pu.lo;
                stage[0][k][2 * q + 1] = pu.hi;
                stage[1][k][2 * q] = pv.lo;
                stage[1][k][2 * q + 1] = pv.hi;
            }
        }
        __syncthreads();
        double a[25], b[25];
#pragma unroll
        for (int k = 0; k <= 8; ++k) {
            const int ilo = k > 4 ? k - 4 : 0, ihi = k < 4 ? k : 4;
#pragma unroll
            for (int i = ilo; i <= ihi; ++i) {
                a[i * 5 + (k - i)] = stage[0][k][li + i];
                b[i * 5 + (k - i)] = stage[1][k][li + i];
            }
        }
        mu = median25(a);
        mv = median25(b);
    } else if (do_median) {
        double a[25], b[25];
        int rr[5], cc[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            rr[k] = mirror_edge(r + k - 2, h);
            cc[k] = mirror_edge(c + k - 2, w);
        }
#pragma unroll
        for (int i = 0; i < 5; ++i)
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                size_t q = (size_t)(rr[i] + cc[j]) * hs + rr[i];
                a[i * 5 + j] = du[q];
                b[i * 5 + j] = dv[q];
            }
        mu = median25(a);
        mv = median25(b);
    } else {
        size_t q = (size_t)(r + c) * hs + r;
        mu = du[q];
        mv = dv[q];
    }
    size_t p = (size_t)r * w + c;
    double un = (double)ut[p] + mu;
    double vn = (double)vt[p] + mv;
    if (out_interleaved) {
        size_t o = p * 2;
        if (out_f64) {
            out_f64[o] = un;
            out_f64[o + 1] = vn;
        } else {
            out_f32[o] = (float)un;
            out_f32[o + 1] = (float)vn;
        }
    } else {
        out_f32[p] = (float)un;
        out_f32[(size_t)h * w + p] = (float)vn;
    }
}

__global__ __launch_bounds__(64) void k_median_add(
    const double *__restrict__ duv, size_t fstride, int h, int w, int hs, size_t plane,
    int do_median, const float *__restrict__ u, const float *__restrict__ v, size_t uv_stride,
    float *out_f32, double *out_f64, int out_interleaved, int T)
{
    const int nchunk = hs / 64, ndiag = h + w - 1;
    const unsigned L = xcd_block(blockIdx.x, (unsigned)nchunk * ndiag * T);
    const int t = L / (nchunk * ndiag), rem = L % (nchunk * ndiag);
    const int d = rem / nchunk;
    const size_t o = out_interleaved ? (size_t)t * h * w * 2 : t * uv_stride;
    __shared__ MedStage stage;
    median_add_px(duv + (size_t)t * fstride, h, w, hs, plane, do_median, u + t * uv_stride, v + t * uv_stride,
                  out_f32 ? out_f32 + o : nullptr, out_f64 ? out_f64 + o : nullptr, out_interleaved, d,
                  (rem % nchunk) * 64, stage);
}

hipError_t launch_median_add(const double *duv, size_t frame_stride, const Skew &s, int T,
                             int do_median, const float *u, const float *v,
                             size_t uv_frame_stride, float *out_f32, double *out_f64,
                             int out_interleaved, hipStream_t st)
{
    const size_t nb = (size_t)(s.hs / 64) * (s.h + s.w - 1) * T;
    if (nb > 0x7fffffffu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_median_add, dim3((unsigned)nb), dim3(64), 0, st, duv, frame_stride, s.h,
                       s.w, s.hs, s.plane, do_median, u, v, uv_frame_stride, out_f32, out_f64,
                       out_interleaved, T);
    return hipGetLastError();
}

// Median of row-major planes (stand-alone entry point), via a copy.
__global__ void k_median_plane(const double *__restrict__ src, double *__restrict__ dst, int h,
                               int w)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, n = blockIdx.z;
    if (c >= w) return;
    const double *f = src + (size_t)n * h * w;
    double a[25];
    int rr[5], cc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        rr[k] = mirror_edge(r + k - 2, h);
        cc[k] = mirror_edge(c + k - 2, w);
    }
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) a[i * 5 + j] = f[(size_t)rr[i] * w + cc[j]];
    dst[((size_t)n * h + r) * w + c] = median25(a);
}

hipError_t launch_median_plane(double *f, int N, int h, int w, double *tmp, hipStream_t st)
{
    hipError_t e = hipMemcpyAsync(tmp, f, sizeof(double) * (size_t)N * h * w,
                                  hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_median_plane, dim3((w + 63) / 64, h, N), dim3(64), 0, st, tmp, f, h, w);
    return hipGetLastError();
}

// ------------------------------------------------- final cubic upsampling
// cv2.resize(w, (n, m), interpolation=INTER_CUBIC) on the float64 flow when
// min_level > 0 (core/optical_flow.py:615-616), restated from OpenCV's
// public behaviour: source coordinate computed in float, Keys cubic
// coefficients (A = -0.75) in float, replicated borders, horizontal pass then
// vertical pass, float64 sums.
__global__ void k_cv_resize_h(const double *__restrict__ src, int T, int h, int w, int C,
                              double *__restrict__ tmp, int ow, double scale_x)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    if (x >= ow) return;
    float f = (float)((x + 0.5) * scale_x - 0.5);
    int ix = (int)floorf(f);
    float cf[4];
    cv_cubic_coeffs(f - (float)ix, cf);
    const double *row = src + ((size_t)t * h + y) * w * C;
    for (int c = 0; c < C; ++c) {
        double s = 0;
        for (int k = 0; k < 4; ++k) s += row[(size_t)clampi(ix - 1 + k, 0, w - 1) * C + c] * (double)cf[k];
        tmp[(((size_t)t * h + y) * ow + x) * C + c] = s;
    }
}

__global__ void k_cv_resize_v(const double *__restrict__ tmp, int T, int h, int ow, int C,
                              double *__restrict__ dst, int oh, double scale_y)
{
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    if (x >= ow) return;
    float f = (float)((y + 0.5) * scale_y - 0.5);
    int iy = (int)floorf(f);
    float cf[4];
    cv_cubic_coeffs(f - (float)iy, cf);
    for (int c = 0; c < C; ++c) {
        double s = 0;
        for (int k = 0; k < 4; ++k)
            s += tmp[(((size_t)t * h + clampi(iy - 1 + k, 0, h - 1)) * ow + x) * C + c] * (double)cf[k];
        dst[(((size_t)t * oh + y) * ow + x) * C + c] = s;
    }
}

hipError_t launch_cv_resize(const double *src, int T, int h, int w, int C, double *tmp,
                            double *dst, int oh, int ow, hipStream_t st)
{
    double sx = 1.0 / ((double)ow / w), sy = 1.0 / ((double)oh / h);
    hipLaunchKernelGGL(k_cv_resize_h, dim3((ow + 63) / 64, h, T), dim3(64), 0, st, src, T, h, w, C, tmp,
                       ow, sx);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cv_resize_v, dim3((ow + 63) / 64, oh, T), dim3(64), 0, st, tmp, T, h, ow, C, dst,
                       oh, sy);
    return hipGetLastError();
}

// ---------------------------------------------------------------- helpers
__global__ void k_fill_f64(double *p, size_t n, double v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t st)
{
    size_t blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

__global__ void k_clear_flags(int *flags, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[(size_t)i * kChainFlagStride] = 0;
}

hipError_t launch_clear_flags(int *flags, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clear_flags, dim3((n + 255) / 256), dim3(256), 0, st, flags, n);
    return hipGetLastError();
}

__global__ void k_broadcast_f32(const float *src, size_t n, float *dst, size_t dst_stride)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    int t = blockIdx.y;
    if (i < n) dst[t * dst_stride + i] = src[i];
}

hipError_t launch_broadcast_f32(const float *src, size_t n, float *dst, size_t dst_stride, int T,
                                hipStream_t st)
{
    hipLaunchKernelGGL(k_broadcast_f32, dim3((unsigned)((n + 255) / 256), T), dim3(256), 0, st, src,
                       n, dst, dst_stride);
    return hipGetLastError();
}

// Row-major float32 planes (N, h, w) -> skewed float64 planes.
__global__ void k_to_skew_f32(const float *__restrict__ src, int h, int w, int hs, size_t plane,
                              double *__restrict__ dst)
{
    int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x, n = blockIdx.y;
    int c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    dst[(size_t)n * plane + (size_t)d * hs + r] = (double)src[((size_t)n * h + r) * w + c];
}

hipError_t launch_to_skew_f32(const float *src, int N, const Skew &s, double *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_skew_f32, dim3(s.h + s.w - 1, N, (s.hs + 1023) / 1024),
                       dim3(s.hs < 1024 ? s.hs : 1024), 0, st, src, s.h, s.w, s.hs, s.plane, dst);
    return hipGetLastError();
}

// Row-major float32 planes (N, h, w) -> skewed float32 planes.
__global__ void k_to_skew_f32f(const float *__restrict__ src, int h, int w, int hs, size_t plane,
                               float *__restrict__ dst)
{
    int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x, n = blockIdx.y;
    int c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    dst[(size_t)n * plane + (size_t)d * hs + r] = src[((size_t)n * h + r) * w + c];
}

hipError_t launch_to_skew_f32f(const float *src, int N, const Skew &s, float *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_to_skew_f32f, dim3(s.h + s.w - 1, N, (s.hs + 1023) / 1024),
                       dim3(s.hs < 1024 ? s.hs : 1024), 0, st, src, s.h, s.w, s.hs, s.plane, dst);
    return hipGetLastError();
}

// (m, n, K) interleaved padded float64 -> K skewed planes of the interior.
__global__ void k_padded_to_skew_f64(const double *__restrict__ src, int m, int n, int K, int hs,
                                     size_t plane, double *__restrict__ dst)
{
    int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x, k = blockIdx.y;
    int h = m - 2, w = n - 2, c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    dst[(size_t)k * plane + (size_t)d * hs + r] = src[((size_t)(r + 1) * n + (c + 1)) * K + k];
}

hipError_t launch_padded_to_skew_f64(const double *src, int m, int n, int K, const Skew &s,
                                     double *dst, hipStream_t st)
{
    hipLaunchKernelGGL(k_padded_to_skew_f64, dim3(s.h + s.w - 1, K, (s.hs + 1023) / 1024),
                       dim3(s.hs < 1024 ? s.hs : 1024), 0, st, src, m, n, K, s.hs, s.plane, dst);
    return hipGetLastError();
}

// Skewed du/dv -> padded (m, n) row-major; the ring is the reference's last
// set_boundary_2d copy (rows first, then columns, so corners take the
// diagonal interior neighbour), taken from the start-of-last-sweep values.
__global__ void k_skew_to_padded_f64(const double *__restrict__ duv,
                                     const double *__restrict__ ring, int h, int w, int hs,
                                     size_t plane, double *du, double *dv)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y;
    int n = w + 2;
    if (i >= n) return;
    bool border = (j == 0 || j == h + 1 || i == 0 || i == w + 1);
    int r = clampi(j - 1, 0, h - 1), c = clampi(i - 1, 0, w - 1);
    size_t q = (size_t)(r + c) * hs + r;
    const double *src = border ? ring : duv;
    du[(size_t)j * n + i] = src[q];
    dv[(size_t)j * n + i] = src[plane + q];
}

hipError_t launch_skew_to_padded_f64(const double *duv, const double *ring, const Skew &s,
                                     double *du, double *dv, hipStream_t st)
{
    hipLaunchKernelGGL(k_skew_to_padded_f64, dim3((s.w + 2 + 63) / 64, s.h + 2), dim3(64), 0, st,
                       duv, ring, s.h, s.w, s.hs, s.plane, du, dv);
    return hipGetLastError();
}

// cu, cv from padded float64 u, v (stand-alone level solver entry point):
// same sum as smooth_rhs, reading the reference's edge-padded arrays.
__global__ void k_cuv_padded(const double *__restrict__ u, const double *__restrict__ v, int m,
                             int n, double ax, double ay, int hs, size_t plane,
                             double *__restrict__ cuv)
{
    int d = blockIdx.x, r = blockIdx.z * blockDim.x + threadIdx.x;
    int h = m - 2, w = n - 2, c = d - r;
    if (r >= h || c < 0 || c >= w) return;
    size_t cc = (size_t)(r + 1) * n + (c + 1);
    size_t nb[4] = {cc - 1, cc + 1, cc + n, cc - n};
    double wgt[4] = {ax, ax, ay, ay};
    double su = wgt[0] * (u[nb[0]] - u[cc]), sv = wgt[0] * (v[nb[0]] - v[cc]);
    for (int k = 1; k < 4; ++k) {
        su = su + wgt[k] * (u[nb[k]] - u[cc]);
        sv = sv + wgt[k] * (v[nb[k]] - v[cc]);
    }
    size_t q = (size_t)d * hs + r;
    cuv[q] = su;
    cuv[plane + q] = sv;
}

hipError_t launch_cuv_padded(const double *u, const double *v, int m, int n, double ax,
                             double ay, const Skew &s, double *cuv, hipStream_t st)
{
    hipLaunchKernelGGL(k_cuv_padded, dim3(s.h + s.w - 1, 1, (s.hs + 1023) / 1024), dim3(s.hs < 1024 ? s.hs : 1024),
                       0, st, u, v, m, n, ax, ay, s.hs, s.plane, cuv);
    return hipGetLastError();
}

// ------------------------------------------------------------- work lists
// One launch per stage over the frames of every group of the level step
// (descriptors gs, item prefix sums s).  Each kernel forms its frames'
// values with the same device functions, in the same order, as the
// launches of one group's level loop.
__device__ __forceinline__ int wl_group(const WlStarts &s, int item)
{
    int g = 0;
#pragma unroll
    for (int k = 1; k < kWlMaxGroups; ++k)
        if (k < s.n && item >= s.start[k]) g = k;
    return g;
}

// level images of the moving and fixed frames (k_pyr_v)
__global__ __launch_bounds__(64) void k_wl_pyr_v(const WlGroup *__restrict__ gs, WlStarts s)
{
    const int item = blockIdx.z, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= G.w || y >= G.h) return;
    const int t = item - s.start[gi], P = G.P_img;
    const int32_t *iy = G.img_iy + (size_t)y * P;
    const float *wy = G.img_wy + (size_t)y * P;
    for (int c = 0; c < G.C; ++c) {
        const size_t n = (size_t)t * G.C + c, o = (n * G.h + y) * G.w + x;
        const float *col = G.pyr + n * G.H * G.Wsum + G.xoff + x;
        G.f2l[o] = (float)tap_sum(P, iy, wy, [&](int i) { return col[(size_t)i * G.Wsum]; });
        const float *col1 = G.pyr1 + n * G.H * G.Wsum + G.xoff + x;
        G.f1l[o] = (float)tap_sum(P, iy, wy, [&](int i) { return col1[(size_t)i * G.Wsum]; });
    }
}

hipError_t launch_wl_pyr_v(const WlGroup *g, const WlStarts &s, int max_oh, int max_ow, hipStream_t st)
{
    hipLaunchKernelGGL(k_wl_pyr_v, dim3((max_ow + 63) / 64, max_oh, s.start[s.n]), dim3(64), 0, st, g, s);
    return hipGetLastError();
}

// the level's initial flow: the initial flow at the coarsest level
// (k_broadcast_f32), the previous level's flow upsampled below it
// (k_resize_fused<4>)
__global__ __launch_bounds__(64) void k_wl_flow(const WlGroup *__restrict__ gs, WlStarts s)
{
    const int item = blockIdx.z, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= G.w || y >= G.h) return;
    const int t = item - s.start[gi];
    const size_t hw = (size_t)G.h * G.w, p = (size_t)y * G.w + x;
    float *dst = G.uvl + (size_t)t * 2 * hw;
    if (G.li == 0) {
        dst[p] = G.uvtop[p];
        dst[hw + p] = G.uvtop[hw + p];
    } else {
        const size_t phw = (size_t)G.ph * G.pw;
        for (int k = 0; k < 2; ++k)
            dst[k * hw + p] = resize_fused_px<4>(G.uvn + ((size_t)t * 2 + k) * phw, G.pw, G.up_ix, G.up_wx,
                                                 G.up_iy, G.up_wy, x, y);
    }
}

hipError_t launch_wl_flow(const WlGroup *g, const WlStarts &s, int max_h, int max_w, hipStream_t st)
{
    hipLaunchKernelGGL(k_wl_flow, dim3((max_w + 63) / 64, max_h, s.start[s.n]), dim3(64), 0, st, g, s);
    return hipGetLastError();
}

// warp of the moving level images below the coarsest level (k_warp, the
// 64F path of the float64 level images)
__global__ __launch_bounds__(64) void k_wl_warp(const WlGroup *__restrict__ gs, WlStarts s)
{
    const int item = blockIdx.z, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (G.li == 0 || x >= G.w || y >= G.h) return;
    const int t = item - s.start[gi];
    const size_t hw = (size_t)G.h * G.w, p = (size_t)y * G.w + x, fo = (size_t)t * G.C * hw;
    const float *u = G.uvl + (size_t)t * 2 * hw;
    warp_cubic_px<float, float, double>(G.f2l + fo, G.f1l + fo, G.f2w + fo, G.C, G.h, G.w, x, y, u[p], u[hw + p],
                                        G.h_col, G.h_row, 0, 0.f, 0.f);
}

hipError_t launch_wl_warp(const WlGroup *g, const WlStarts &s, int max_h, int max_w, hipStream_t st)
{
    hipLaunchKernelGGL(k_wl_warp, dim3((max_w + 63) / 64, max_h, s.start[s.n]), dim3(64), 0, st, g, s);
    return hipGetLastError();
}

// motion tensor and cu, cv (k_tensor); the warped image is float32 below the
// coarsest level, the grid spacing is unit at full resolution
template <int KIND>
__global__ __launch_bounds__(kTensorCols * kTensorRows, KIND == kTensorCS ? 2 : 6) void k_wl_tensor(const WlGroup *__restrict__ gs,
                                                                          WlStarts s, double eps)
{
    __shared__ double stage[6][kTensorRows][kTensorCols];
    const int item = blockIdx.y, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int h = G.h, w = G.w, C = G.C;
    const int tiles_x = (w + kTensorCols - 1) / kTensorCols, tiles_y = (h + kTensorRows - 1) / kTensorRows;
    if ((int)blockIdx.x >= tiles_x * tiles_y) return;
    const int r0 = (blockIdx.x / tiles_x) * kTensorRows, c0 = (blockIdx.x % tiles_x) * kTensorCols;
    const int t = item - s.start[gi];
    const size_t hw = (size_t)h * w;
    const float *f1t = G.f1l + (size_t)t * C * hw;
    const float *f2t = (G.li == 0 ? G.f2l : G.f2w) + (size_t)t * C * hw;
    const float *ut = G.uvl + (size_t)t * 2 * hw;
    double *Jt = G.work + (size_t)t * G.fs, *ct = Jt + (size_t)6 * C * G.plane;
    const bool unit = G.h_row == 1.0 && G.h_col == 1.0;
#define FRH_WLT(P2, U)                                                                                    \
    tensor_tile<KIND, P2, U>(stage, f1t, f2t, C, h, w, G.hs, G.plane, G.h_row, G.h_col, eps, ut, ut + hw, G.axw, \
                             G.ayw, Jt, ct, r0, c0)
    if (G.li == 0) {
        if (unit) FRH_WLT(false, true);
        else FRH_WLT(false, false);
    } else {
        if (unit) FRH_WLT(true, true);
        else FRH_WLT(true, false);
    }
#undef FRH_WLT
}

hipError_t launch_wl_tensor(int kind, const WlGroup *g, const WlStarts &s, int max_tiles, double eps,
                            hipStream_t st)
{
    const dim3 grid(max_tiles, s.start[s.n]), b(kTensorCols, kTensorRows);
    switch (kind) {
    case kTensorGC: hipLaunchKernelGGL(k_wl_tensor<kTensorGC>, grid, b, 0, st, g, s, eps); break;
    case kTensorGray: hipLaunchKernelGGL(k_wl_tensor<kTensorGray>, grid, b, 0, st, g, s, eps); break;
    case kTensorCS: hipLaunchKernelGGL(k_wl_tensor<kTensorCS>, grid, b, 0, st, g, s, eps); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// lag-block coefficients (k_assemble with the float32 weight planes); one
// work-group per 256 rows of a diagonal, small enough to be placed on CUs
// that the other form's sweeps occupy in part
constexpr int kWlAsmRows = 256;

__global__ __launch_bounds__(kWlAsmRows) void k_wl_assemble(const WlGroup *__restrict__ gs, WlStarts s,
                                                            int psi_one, int zero_init)
{
    const int item = blockIdx.y, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int d = blockIdx.x, r = blockIdx.z * kWlAsmRows + threadIdx.x, c = d - r;
    if (r >= G.h || c < 0 || c >= G.w) return;
    const size_t pl = G.plane, q = (size_t)d * G.hs + r;
    const double *Jt = G.work + (size_t)(item - s.start[gi]) * G.fs;
    const double *ct = Jt + (size_t)6 * G.C * pl;
    double *co = const_cast<double *>(ct) + 2 * pl;
    const double *dt = co + 5 * pl;
#define FRH_WLA(CC) assemble_cell<CC, float>(Jt, G.Wlev, ct, dt, co, pl, q, G.axw, G.ayw, G.ad, psi_one, zero_init)
    switch (G.C) {
    case 1: FRH_WLA(1); break;
    case 2: FRH_WLA(2); break;
    case 3: FRH_WLA(3); break;
    case 4: FRH_WLA(4); break;
    default: break;
    }
#undef FRH_WLA
}

hipError_t launch_wl_assemble(const WlGroup *g, const WlStarts &s, int max_diag, int max_hs, int psi_one,
                              int zero_init, hipStream_t st)
{
    if (max_hs < 64 || max_hs > 1024) return hipErrorInvalidValue;
    const int bx = max_hs < kWlAsmRows ? max_hs : kWlAsmRows;
    hipLaunchKernelGGL(k_wl_assemble, dim3(max_diag, s.start[s.n], (max_hs + bx - 1) / bx), dim3(bx), 0, st, g, s,
                       psi_one, zero_init);
    return hipGetLastError();
}

// one wavefront pass per frame (k_sweep); the work-group's waves past the
// frame's rows (hs / 64 waves) leave at once, and the barriers of the pass
// wait for the waves that remain
template <int P, int MAXT>
__global__ __launch_bounds__(MAXT) void k_wl_sweep(const WlGroup *__restrict__ gs, WlStarts s, int zero_init)
{
    __shared__ double xb[2][MAXT / 64][2][P + 1][2];
    const int item = blockIdx.x, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int nthr = G.hs;
    if ((int)threadIdx.x >= nthr) return;
    SweepArgs a{};
    a.coef = G.work + (size_t)(6 * G.C + 2) * G.plane;
    a.duv = const_cast<double *>(a.coef) + 5 * G.plane;
    a.ring = nullptr;
    a.fstride = G.fs;
    a.plane = G.plane;
    a.h = G.h; a.w = G.w; a.hs = G.hs;
    a.ax = G.axw; a.ay = G.ayw;
    a.zero_init = zero_init;
    a.r0 = 0;
    a.hlev = G.h;
    sweep_frame<P, MAXT, 0, false>(a, item - s.start[gi], nthr, xb);
}

hipError_t launch_wl_sweep(int P, int form1024, const WlGroup *g, const WlStarts &s, int zero_init,
                           hipStream_t st)
{
    const dim3 grid(s.start[s.n]);
    if (form1024) {
        switch (P) {
        case 1: hipLaunchKernelGGL((k_wl_sweep<1, 1024>), grid, dim3(1024), 0, st, g, s, zero_init); break;
        case 2: hipLaunchKernelGGL((k_wl_sweep<2, 1024>), grid, dim3(1024), 0, st, g, s, zero_init); break;
        default: return hipErrorInvalidValue;
        }
    } else {
        switch (P) {
        case 1: hipLaunchKernelGGL((k_wl_sweep<1, 512>), grid, dim3(512), 0, st, g, s, zero_init); break;
        case 2: hipLaunchKernelGGL((k_wl_sweep<2, 512>), grid, dim3(512), 0, st, g, s, zero_init); break;
        case 3: hipLaunchKernelGGL((k_wl_sweep<3, 512>), grid, dim3(512), 0, st, g, s, zero_init); break;
        case 4: hipLaunchKernelGGL((k_wl_sweep<4, 512>), grid, dim3(512), 0, st, g, s, zero_init); break;
        case 5: hipLaunchKernelGGL((k_wl_sweep<5, 512>), grid, dim3(512), 0, st, g, s, zero_init); break;
        default: return hipErrorInvalidValue;
        }
    }
    return hipGetLastError();
}

// median + u += du (k_median_add): next level's flow planes, the output
// (float32 / float64, (h, w, 2) per frame) or the coarse float64 flow
__global__ __launch_bounds__(64) void k_wl_median(const WlGroup *__restrict__ gs, WlStarts s)
{
    const int item = blockIdx.y, gi = wl_group(s, item);
    const WlGroup &G = gs[gi];
    const int nchunk = G.hs / 64, L = blockIdx.x;
    if (L >= nchunk * (G.h + G.w - 1)) return;
    const int t = item - s.start[gi];
    const size_t hw = (size_t)G.h * G.w;
    const double *du = G.work + (size_t)t * G.fs + (size_t)(6 * G.C + 7) * G.plane;
    const float *ut = G.uvl + (size_t)t * 2 * hw;
    float *o32 = nullptr;
    double *o64 = nullptr;
    if (G.out_kind == 0) o32 = G.uvn + (size_t)t * 2 * hw;
    else if (G.out_kind == 1) o32 = (float *)G.out + (size_t)t * 2 * hw;
    else o64 = (double *)G.out + (size_t)t * 2 * hw;
    __shared__ MedStage stage;
    median_add_px(du, G.h, G.w, G.hs, G.plane, G.do_median, ut, ut + hw, o32, o64, G.out_kind != 0, L / nchunk,
                  (L % nchunk) * 64, stage);
}

hipError_t launch_wl_median(const WlGroup *g, const WlStarts &s, int max_blocks, hipStream_t st)
{
    hipLaunchKernelGGL(k_wl_median, dim3(max_blocks, s.start[s.n]), dim3(64), 0, st, g, s);
    return hipGetLastError();
}

}  // namespace frh
