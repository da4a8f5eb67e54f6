// Differentiable single-level registration: forward and hand-written
// backward kernels for get_displacement(..., levels=0) followed by a
// continuous bicubic warp (core/optical_flow.py:476-617 with one level,
// core/level_solver.py:158-369, core/warping.py:160-234 without the 1/32-px
// map rounding).  Everything is float64 and laid out row-major per frame:
//   images (T, C, h, w); states x (T, 2, h, w) = (du, dv); J (T, 6, C, h, w);
//   coefficients (T, 6, h, w) = (A12, Bu, Bv, wu = omega/Du, wv = omega/Dv, b).
// The SOR adjoint is the exact reverse-mode derivative of one column-major
// lexicographic sweep, executed as a reverse diagonal wavefront; the
// coefficient and psi-refresh adjoint is written out by hand, and the
// motion-tensor derivative comes from forward-mode dual numbers over each
// cell's 3x3 moving-frame neighbourhood.

#include <hip/hip_runtime.h>

#include <cmath>
#include <type_traits>
#include <utility>

#include "../../include/flowreg_hip.h"
#include "frh_internal.h"
#include "frh_device.h"
#include "frh_median.h"

namespace frh {
namespace grad {

constexpr double kOmega = 1.95;

// ------------------------------------------------------------ dual numbers
template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ inline Dual<N> dconst(double v)
{
    Dual<N> r;
    r.v = v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = 0.0;
    return r;
}

template <int N>
__device__ inline Dual<N> dvar(double v, int i)
{
    Dual<N> r = dconst<N>(v);
    r.d[i] = 1.0;
    return r;
}

#define FRH_DUAL_BINOP(OP, VAL, DER)                                        \
    template <int N>                                                        \
    __device__ inline Dual<N> operator OP(const Dual<N> &a, const Dual<N> &b) \
    {                                                                       \
        Dual<N> r;                                                          \
        r.v = VAL;                                                          \
        _Pragma("unroll") for (int i = 0; i < N; ++i) r.d[i] = DER;         \
        return r;                                                           \
    }
FRH_DUAL_BINOP(+, a.v + b.v, a.d[i] + b.d[i])
FRH_DUAL_BINOP(-, a.v - b.v, a.d[i] - b.d[i])
FRH_DUAL_BINOP(*, a.v * b.v, a.d[i] * b.v + a.v * b.d[i])
FRH_DUAL_BINOP(/, a.v / b.v, (a.d[i] * b.v - a.v * b.d[i]) / (b.v * b.v))
#undef FRH_DUAL_BINOP

template <int N>
__device__ inline Dual<N> operator*(double s, const Dual<N> &a)
{
    Dual<N> r;
    r.v = s * a.v;
#pragma unroll
    for (int i = 0; i < N; ++i) r.d[i] = s * a.d[i];
    return r;
}

template <int N>
__device__ inline Dual<N> operator*(const Dual<N> &a, double s)
{
    return s * a;
}

template <int N>
__device__ inline Dual<N> operator+(const Dual<N> &a, double s)
{
    Dual<N> r = a;
    r.v += s;
    return r;
}

template <int N>
__device__ inline Dual<N> operator+(double s, const Dual<N> &a)
{
    return a + s;
}

template <int N>
__device__ inline Dual<N> operator-(const Dual<N> &a, double s)
{
    return a + (-s);
}

template <typename R>
__device__ inline R rconst(double v)
{
    if constexpr (std::is_same<R, double>::value)
        return v;
    else
        return dconst<sizeof(R) / sizeof(double) - 1>(v);
}

// ---------------------------------------------------------- motion tensor
// core/optical_flow.py:73-160 on one cell from its clamped 3x3
// neighbourhoods (row-major, centre 4) of f1 and f2, generic in f2's type.
template <typename R>
__device__ inline void gc_cell(const double *a, const R *b, double hy, double hx, R *J)
{
    auto gx = [&](int row) -> R {  // 0.5 (d/dx f1 + d/dx f2) at neighbourhood row
        const double g1 = (a[row * 3 + 2] - a[row * 3 + 0]) / (2.0 * hx);
        const R g2 = (1.0 / (2.0 * hx)) * (b[row * 3 + 2] - b[row * 3 + 0]);
        return 0.5 * (g2 + g1);
    };
    auto ft = [&](int k) -> R { return b[k] + (-a[k]); };
    const R fxy = (1.0 / (2.0 * hy)) * (gx(2) - gx(0));
    const R fxt = (1.0 / (2.0 * hx)) * (ft(5) - ft(3));
    const R fyt = (1.0 / (2.0 * hy)) * (ft(7) - ft(1));
    const double fxx1 = (a[3] - 2 * a[4] + a[5]) / (hx * hx);
    const double fyy1 = (a[1] - 2 * a[4] + a[7]) / (hy * hy);
    const R fxx2 = (1.0 / (hx * hx)) * ((b[3] - 2.0 * b[4]) + b[5]);
    const R fyy2 = (1.0 / (hy * hy)) * ((b[1] - 2.0 * b[4]) + b[7]);
    const R fxx = 0.5 * (fxx2 + fxx1);
    const R fyy = 0.5 * (fyy2 + fyy1);
    const R one = rconst<R>(1.0);
    const R regx = one / ((fxx * fxx + fxy * fxy) + 1e-6);
    const R regy = one / ((fxy * fxy + fyy * fyy) + 1e-6);
    J[0] = regx * (fxx * fxx) + regy * (fxy * fxy);
    J[1] = regx * (fxy * fxy) + regy * (fyy * fyy);
    J[2] = regx * (fxt * fxt) + regy * (fyt * fyt);
    J[3] = regx * fxx * fxy + regy * fxy * fyy;
    J[4] = regx * fxx * fxt + regy * fxy * fyt;
    J[5] = regx * fxy * fxt + regy * fyy * fyt;
}

__device__ inline void gather3(const double *f, int h, int w, int r, int c, double *out, int *idx)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            int rr = r + i - 1, cc = c + j - 1;
            rr = rr < 0 ? 0 : (rr > h - 1 ? h - 1 : rr);
            cc = cc < 0 ? 0 : (cc > w - 1 ? w - 1 : cc);
            idx[i * 3 + j] = rr * w + cc;
            out[i * 3 + j] = f[rr * w + cc];
        }
}

// J (T, 6, C, h, w) from fixed (C, h, w) and moving (T, C, h, w)
__global__ __launch_bounds__(64) void k_gtensor_fwd(const double *f1, const double *f2, int T, int C, int h, int w,
                              double hy, double hx, double *J)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, tk = blockIdx.z;
    if (c >= w) return;
    const int t = tk / C, k = tk % C;
    const size_t hw = (size_t)h * w;
    double a[9], b[9], Jc[6];
    int idx[9];
    gather3(f1 + k * hw, h, w, r, c, a, idx);
    gather3(f2 + ((size_t)t * C + k) * hw, h, w, r, c, b, idx);
    gc_cell<double>(a, b, hy, hx, Jc);
    for (int e = 0; e < 6; ++e) J[(((size_t)t * 6 + e) * C + k) * hw + (size_t)r * w + c] = Jc[e];
}

// dL/dmoving += sum_e gJ_e dJ_e/dmoving over the 3x3 inputs of each cell
__global__ __launch_bounds__(64) void k_gtensor_bwd(const double *f1, const double *f2, const double *gJ, int T, int C,
                              int h, int w, double hy, double hx, double *gf2)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, tk = blockIdx.z;
    if (c >= w) return;
    const int t = tk / C, k = tk % C;
    const size_t hw = (size_t)h * w;
    double a[9], bv[9];
    int idx[9];
    gather3(f1 + k * hw, h, w, r, c, a, idx);
    gather3(f2 + ((size_t)t * C + k) * hw, h, w, r, c, bv, idx);
    Dual<9> b[9], Jc[6];
    for (int i = 0; i < 9; ++i) b[i] = dvar<9>(bv[i], i);
    gc_cell<Dual<9>>(a, b, hy, hx, Jc);
    double g[6];
    for (int e = 0; e < 6; ++e) g[e] = gJ[(((size_t)t * 6 + e) * C + k) * hw + (size_t)r * w + c];
    double *out = gf2 + ((size_t)t * C + k) * hw;
    for (int i = 0; i < 9; ++i) {
        double s = 0.0;
        for (int e = 0; e < 6; ++e) s += g[e] * Jc[e].d[i];
        if (s != 0.0) atomicAdd(out + idx[i], s);
    }
}

// -------------------------------------------------------- coefficients
// Lag-block coefficients of core/level_solver.py:229-255,317-353 for one
// cell, generic in the scalar type: inputs the cell's state (x, y) at the
// refresh, ax, ay, a_k, J (6C), W (C) and the smoothness terms (cu, cv) of
// the level's initial flow (a_smooth = 1).
constexpr int kMaxC = 2;

__device__ inline double psi_of(double E, double a)  // a (max(E, 0) + 1e-5) ** (a - 1)
{
    const double e = E < 0 ? 0.0 : E;
    return a * pow(e + 1e-5, a - 1.0);
}

__device__ inline void coef_cell(double x, double y, double ax, double ay, const double *a, const double *J,
                                 const double *W, int C, int psi_one, double nlr, double nud, double cu,
                                 double cv, double *K)
{
    const double ws = (ax + ax) + (ay + ay);
    double Du = ws, Dv = ws, A12 = 0.0, Bu = cu, Bv = cv;
    for (int k = 0; k < C; ++k) {
        const double j11 = J[0 * C + k], j22 = J[1 * C + k], j33 = J[2 * C + k];
        const double j12 = J[3 * C + k], j13 = J[4 * C + k], j23 = J[5 * C + k];
        double psi = 1.0;
        if (!psi_one)
            psi = psi_of(j11 * x * x + j22 * y * y + 2.0 * (j23 * y) + 2.0 * (j12 * x * y) + 2.0 * (j13 * x) + j33,
                         a[k]);
        const double wp = W[k] * psi;
        Du += wp * j11;
        Dv += wp * j22;
        A12 += wp * j12;
        Bu -= wp * j13;
        Bv -= wp * j23;
    }
    K[0] = A12;
    K[1] = Bu;
    K[2] = Bv;
    K[3] = kOmega / Du;
    K[4] = kOmega / Dv;
    K[5] = nlr * ax + nud * ay;
}

__device__ inline void edge_counts(int h, int w, int r, int c, double &nlr, double &nud)
{
    nlr = (double)((c == 0) + (c == w - 1));
    nud = (double)((r == 0) + (r == h - 1));
}

// Neighbour sums of the edge-padded initial flow f at cell (r, c):
// sx = sum over in-level left/right neighbours of (f_nb - f_c), sy likewise
// up/down (neighbours in the padding ring equal f_c and add nothing).
__device__ inline void flow_diffs(const double *f, int h, int w, int r, int c, double &sx, double &sy)
{
    const size_t p = (size_t)r * w + c;
    const double fc = f[p];
    sx = (c > 0 ? f[p - 1] - fc : 0.0) + (c < w - 1 ? f[p + 1] - fc : 0.0);
    sy = (r > 0 ? f[p - w] - fc : 0.0) + (r < h - 1 ? f[p + w] - fc : 0.0);
}

// Neighbour sums of a skewed plane f (one value per `stride` doubles: a
// component of cell-interleaved records) at cell (r, c) in slot q: the same
// operator as flow_diffs ((r, c +- 1) sit at q +- hs, (r +- 1, c) at
// q +- (hs + 1)).
__device__ inline void flow_diffs_skew(const double *f, int stride, int h, int w, int hs, int r, int c, size_t q,
                                       double &sx, double &sy)
{
    const double fc = f[q * stride];
    sx = (c > 0 ? f[(q - hs) * stride] - fc : 0.0) + (c < w - 1 ? f[(q + hs) * stride] - fc : 0.0);
    sy = (r > 0 ? f[(q - hs - 1) * stride] - fc : 0.0) + (r < h - 1 ? f[(q + hs + 1) * stride] - fc : 0.0);
}

// The solve's state, coefficients and their adjoints live in the skewed
// layout of the production solver: cell (r, c) of a frame's h x w level at
// slot (r + c) * hs + r of a plane of (h + w - 1) * hs slots, hs = 64 *
// ceil(h / 64), so that one anti-diagonal is one contiguous run.  A slot
// holds a cell's whole record, interleaved: the (u, v) pair of a state or
// state adjoint (16 bytes), the 6 coefficients or coefficient adjoints (48
// bytes), so a lane reads a cell's record with one or three 16-byte loads
// (the sweeps are bound by their vector memory instructions: 0.58 against
// 0.90 ms per forward sweep at 32 frames of 512 x 512 with one plane per
// component).  Padding slots are never written by the solve's kernels.
struct GSkew {
    int h, w, hs;
    size_t plane;
};

constexpr unsigned kPairB = 16, kCoefB = 48;  // bytes of a state / coefficient record

// byte offset of the record of the slot at byte offset q8 of a plane of
// doubles (slot_offset), or kOutOfRange
__device__ inline unsigned rec_offset(unsigned q8, unsigned rec)
{
    return q8 == kOutOfRange ? kOutOfRange : q8 / 8u * rec;
}

__device__ inline F64x2 load_pair(__amdgpu_buffer_rsrc_t r, unsigned q8)
{
    return buf_load_f64x2(r, rec_offset(q8, kPairB), 0);
}

__device__ inline void load_coef6(__amdgpu_buffer_rsrc_t r, unsigned q8, double *K)
{
    const unsigned o = rec_offset(q8, kCoefB);
    const F64x2 a = buf_load_f64x2(r, o, 0), b = buf_load_f64x2(r, o, 16), c = buf_load_f64x2(r, o, 32);
    K[0] = a.lo;
    K[1] = a.hi;
    K[2] = b.lo;
    K[3] = b.hi;
    K[4] = c.lo;
    K[5] = c.hi;
}

__device__ inline void store_pair(double u, double v, __amdgpu_buffer_rsrc_t r, unsigned q8)
{
    buf_store_f64x2(F64x2{u, v}, r, rec_offset(q8, kPairB), 0);
}

__device__ inline void store_coef6(const double *K, __amdgpu_buffer_rsrc_t r, unsigned q8)
{
    const unsigned o = rec_offset(q8, kCoefB);
    buf_store_f64x2(F64x2{K[0], K[1]}, r, o, 0);
    buf_store_f64x2(F64x2{K[2], K[3]}, r, o, 16);
    buf_store_f64x2(F64x2{K[4], K[5]}, r, o, 32);
}

__device__ inline bool skew_cell(const GSkew &k, size_t q, int &r, int &c)
{
    r = (int)(q % (size_t)k.hs);
    c = (int)(q / (size_t)k.hs) - r;
    return r < k.h && c >= 0 && c < k.w;
}

// (N, K, h, w) row-major <-> (N, plane, K) skewed with K values per slot
// (padding slots set to zero)
__global__ void k_gskew(const double *src, double *dst, int K, GSkew k, int to_skew)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int n = blockIdx.y;
    if (q >= k.plane) return;
    int r, c;
    const bool ok = skew_cell(k, q, r, c);
    const size_t hw = (size_t)k.h * k.w, p = (size_t)r * k.w + c;
    double *rec = to_skew ? dst + ((size_t)n * k.plane + q) * K : nullptr;
    const double *srec = to_skew ? nullptr : src + ((size_t)n * k.plane + q) * K;
    for (int e = 0; e < K; ++e) {
        const size_t rm = ((size_t)n * K + e) * hw + p;
        if (to_skew)
            rec[e] = ok ? src[rm] : 0.0;
        else if (ok)
            dst[rm] = srec[e];
    }
}

// Cell of a coefficient-kernel thread: a 256-lane work-group covers a tile
// of 16 rows x 16 columns, each wave 4 rows of it, so a wave's row-major
// accesses (J, W, u0 and their adjoints, the most) are 128-byte runs and its
// skewed ones (state, coefficients and their adjoints) 19 anti-diagonal
// runs instead of one cache line per lane (a wave on 1 x 64 cells); the
// work-group's four waves continue each other's anti-diagonal runs, whose
// cache lines they share.
constexpr int kCoefTR = 16, kCoefTC = 16;

__host__ __device__ inline unsigned coef_tiles(int h, int w)
{
    return (unsigned)((w + kCoefTC - 1) / kCoefTC) * (unsigned)((h + kCoefTR - 1) / kCoefTR);
}

__device__ inline bool coef_tile_cell(int h, int w, int &r, int &c)
{
    const int ntx = (w + kCoefTC - 1) / kCoefTC;
    const int tx = (int)blockIdx.x % ntx, ty = (int)blockIdx.x / ntx;
    r = ty * kCoefTR + (int)(threadIdx.x / kCoefTC);
    c = tx * kCoefTC + (int)(threadIdx.x % kCoefTC);
    return r < h && c < w;
}

// x, coef skewed; u0, J, W row-major.
template <int C>
__global__ __launch_bounds__(kCoefTR * kCoefTC) void k_gcoef_fwd(const double *x, const double *u0, const double *J,
                                                                  const double *W, GSkew k, double ax, double ay,
                                                                  const double *a_data, int psi_one, double *coef)
{
    const int t = blockIdx.y;
    const int h = k.h, w = k.w;
    const size_t hw = (size_t)h * w;
    int r, c;
    if (!coef_tile_cell(h, w, r, c)) return;
    const size_t p = (size_t)r * w + c;
    const size_t q = (size_t)(r + c) * k.hs + r;
    double Jr[6 * C], Wr[C], a[C], K[6];
#pragma unroll
    for (int e = 0; e < 6; ++e)
#pragma unroll
        for (int kk = 0; kk < C; ++kk) Jr[e * C + kk] = J[(((size_t)t * 6 + e) * C + kk) * hw + p];
#pragma unroll
    for (int kk = 0; kk < C; ++kk) {
        Wr[kk] = W[kk * hw + p];
        a[kk] = a_data[kk];
    }
    double nlr, nud, cu = 0.0, cv = 0.0;
    edge_counts(h, w, r, c, nlr, nud);
    if (u0) {
        double sx, sy;
        flow_diffs(u0 + (size_t)t * 2 * hw, h, w, r, c, sx, sy);
        cu = ax * sx + ay * sy;
        flow_diffs(u0 + ((size_t)t * 2 + 1) * hw, h, w, r, c, sx, sy);
        cv = ax * sx + ay * sy;
    }
    const double *X = x + ((size_t)t * k.plane + q) * 2;
    coef_cell(X[0], X[1], ax, ay, a, Jr, Wr, C, psi_one, nlr, nud, cu, cv, K);
    double *co = coef + ((size_t)t * k.plane + q) * 6;
#pragma unroll
    for (int e = 0; e < 6; ++e) co[e] = K[e];
}

// Pull-back of the coefficient adjoints gK (T, 6, plane) through coef_cell:
// gx += dK/dx^T gK (T, 2, plane), gJ += dK/dJ^T gK (row-major), per-cell
// partial sums for ax, ay, a_k in gpar (T, 2 + C, plane), and (with u0)
// gu0 += the adjoint of the smoothness terms (row-major): cu = ax * sx(u) +
// ay * sy(u) with the symmetric neighbour-difference operators sx, sy, so
// gu0 = ax * sx(gBu) + ay * sy(gBu).  Reverse mode by hand:
//   gDu = -omega / Du^2 gK3, gDv = -omega / Dv^2 gK4, gws = gDu + gDv
//   per channel: g_wp = gDu j11 + gDv j22 + gK0 j12 - gK1 j13 - gK2 j23,
//   g_psi = W g_wp; psi = a p, p = (e + 1e-5)^(a - 1), e = max(E, 0):
//   dpsi/da = p (1 + a ln(e + 1e-5)), dpsi/dE = a (a - 1) p / (e + 1e-5)
//   (zero for E < 0); E is quadratic in (x, y) and linear in J.
template <int C>
__global__ __launch_bounds__(kCoefTR * kCoefTC) void k_gcoef_bwd(const double *x, const double *u0, const double *J,
                                                                  const double *W, const double *gK, GSkew k,
                                                                  double ax, double ay, const double *a_data,
                                                                  int psi_one, double *gx, double *gu0, double *gJ,
                                                                  double *gpar)
{
    constexpr int NW = kCoefTR * kCoefTC / 64;
    __shared__ double red[NW][2 + C];
    const int t = blockIdx.y;
    const int h = k.h, w = k.w;
    const size_t hw = (size_t)h * w, P = k.plane;
    int r, c;
    const bool in_level = coef_tile_cell(h, w, r, c);
    double gs[2 + C];  // this cell's partials of ax, ay, a_k
#pragma unroll
    for (int i = 0; i < 2 + C; ++i) gs[i] = 0.0;
    if (in_level) {
        const size_t p = (size_t)r * w + c;
        const size_t q = (size_t)(r + c) * k.hs + r;
        const double xv = x[((size_t)t * P + q) * 2], yv = x[((size_t)t * P + q) * 2 + 1];
        double g[6];
#pragma unroll
        for (int e = 0; e < 6; ++e) g[e] = gK[((size_t)t * P + q) * 6 + e];
        // Du, Dv at the refresh (recomputed)
        const double ws = (ax + ax) + (ay + ay);
        double Du = ws, Dv = ws;
        double psi[C], Jc[6 * C], Wc[C];
#pragma unroll
        for (int kk = 0; kk < C; ++kk) {
#pragma unroll
            for (int e = 0; e < 6; ++e) Jc[e * C + kk] = J[(((size_t)t * 6 + e) * C + kk) * hw + p];
            const double j11 = Jc[0 * C + kk], j22 = Jc[1 * C + kk], j33 = Jc[2 * C + kk];
            const double j12 = Jc[3 * C + kk], j13 = Jc[4 * C + kk], j23 = Jc[5 * C + kk];
            psi[kk] = psi_one ? 1.0
                              : psi_of(j11 * xv * xv + j22 * yv * yv + 2.0 * (j23 * yv) + 2.0 * (j12 * xv * yv) +
                                           2.0 * (j13 * xv) + j33,
                                       a_data[kk]);
            Wc[kk] = W[kk * hw + p];
            const double wp = Wc[kk] * psi[kk];
            Du += wp * j11;
            Dv += wp * j22;
        }
        const double gDu = -kOmega / (Du * Du) * g[3], gDv = -kOmega / (Dv * Dv) * g[4];
        const double gws = gDu + gDv;
        double nlr, nud;
        edge_counts(h, w, r, c, nlr, nud);
        double g_ax = 2.0 * gws + nlr * g[5], g_ay = 2.0 * gws + nud * g[5];
        double g_x = 0.0, g_y = 0.0;
#pragma unroll
        for (int kk = 0; kk < C; ++kk) {
            const double j11 = Jc[0 * C + kk], j22 = Jc[1 * C + kk], j33 = Jc[2 * C + kk];
            const double j12 = Jc[3 * C + kk], j13 = Jc[4 * C + kk], j23 = Jc[5 * C + kk];
            const double Wk = Wc[kk], wp = Wk * psi[kk];
            double gj[6];
            gj[0] = gDu * wp;
            gj[1] = gDv * wp;
            gj[2] = 0.0;
            gj[3] = g[0] * wp;
            gj[4] = -g[1] * wp;
            gj[5] = -g[2] * wp;
            if (!psi_one) {
                const double g_psi = Wk * (gDu * j11 + gDv * j22 + g[0] * j12 - g[1] * j13 - g[2] * j23);
                const double E = j11 * xv * xv + j22 * yv * yv + 2.0 * (j23 * yv) + 2.0 * (j12 * xv * yv) +
                                 2.0 * (j13 * xv) + j33;
                const double a = a_data[kk], base = (E < 0 ? 0.0 : E) + 1e-5;
                const double pw = pow(base, a - 1.0);
                gs[2 + kk] = g_psi * pw * (1.0 + a * log(base));
                const double gE = E < 0 ? 0.0 : g_psi * a * (a - 1.0) * pw / base;
                gj[0] += gE * xv * xv;
                gj[1] += gE * yv * yv;
                gj[2] += gE;
                gj[3] += gE * 2.0 * xv * yv;
                gj[4] += gE * 2.0 * xv;
                gj[5] += gE * 2.0 * yv;
                g_x += gE * 2.0 * (j11 * xv + j12 * yv + j13);
                g_y += gE * 2.0 * (j22 * yv + j12 * xv + j23);
            }
#pragma unroll
            for (int e = 0; e < 6; ++e) gJ[(((size_t)t * 6 + e) * C + kk) * hw + p] += gj[e];
        }
        if (u0) {
            const double *gBu = gK + (size_t)t * P * 6 + 1, *gBv = gK + (size_t)t * P * 6 + 2;
            double sx, sy;
            flow_diffs(u0 + (size_t)t * 2 * hw, h, w, r, c, sx, sy);
            g_ax += g[1] * sx;
            g_ay += g[1] * sy;
            flow_diffs(u0 + ((size_t)t * 2 + 1) * hw, h, w, r, c, sx, sy);
            g_ax += g[2] * sx;
            g_ay += g[2] * sy;
            if (gu0) {
                flow_diffs_skew(gBu, 6, h, w, k.hs, r, c, q, sx, sy);
                gu0[((size_t)t * 2) * hw + p] += ax * sx + ay * sy;
                flow_diffs_skew(gBv, 6, h, w, k.hs, r, c, q, sx, sy);
                gu0[((size_t)t * 2 + 1) * hw + p] += ax * sx + ay * sy;
            }
        }
        gx[((size_t)t * P + q) * 2] += g_x;
        gx[((size_t)t * P + q) * 2 + 1] += g_y;
        gs[0] = g_ax;
        gs[1] = g_ay;
    }
    // the tile's sums of the ax, ay, a_k partials (a butterfly in each wave,
    // then the waves in order), one addition per tile to gpar
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < 2 + C; ++i) {
        double v = gs[i];
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
        if (lane == 0) red[wv][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < 2 + C) {
        double v = red[0][threadIdx.x];
#pragma unroll
        for (int j = 1; j < NW; ++j) v += red[j][threadIdx.x];
        gpar[((size_t)t * (2 + C) + threadIdx.x) * P + blockIdx.x] += v;
    }
}


// --------------------------------------------------------------- sweeps
// One column-major lexicographic SOR sweep (core/level_solver.py:257-364),
// one work-group per frame, lane = row, one anti-diagonal d = r + c per
// step:
//   nu = Bu + ax (l + r) + ay (d + u) - A12 ov
//   du' = wu nu + (1 - omega + wu b) ou
//   nv = Bv + ax (lv + rv) + ay (dv + uv) - A12 du'
//   dv' = wv nv + (1 - omega + wv b) ov
// (left / up neighbours new, right / down old; neighbours outside the level
// hold the cell's own start value, folded into b).  The left neighbour is
// the lane's own previous result, the upper one lane r - 1's (a DPP shift
// inside the wave, LDS across waves); the old values and coefficients do not
// depend on the sweep, so they are loaded D steps ahead into a register ring
// with branch-free buffer loads, and the per-step barrier waits for LDS only.
constexpr int kGMaxWaves = 16;

// Ring slot of one diagonal: its coefficients and old values; du, dv (the
// old values of row r + 1) are loaded by lane 63 only -- the other lanes
// take them from lane r + 1's slot of the next diagonal with a DPP shift,
// and the old right neighbour (r, c + 1) is the lane's own slot of the next
// diagonal, so each old value is read from memory once.
struct GFwdSlot {
    double K[6], ou, ov, du, dv;
};

// The sweeps of a lag block as a chain (CHAIN): work-group (stage s, frame
// t) runs sweep s of the block, the same code as a single launch, reading
// stage s - 1's output as it is published.  Stage s writes xs + s sstride;
// with sstride 0 all stages work in place on xs: a stage reads a cell of
// its input (its predecessor's output, published) before it overwrites it,
// and its successor reads the cell only after it is published.
struct GFChain {
    const double *x0;  // stage 0's input
    double *xs;        // stage s's output at xs + s sstride
    size_t sstride;
    int *flags;        // stage s, frame t: flags[(s Tpad + t) kChainFlagStride]
    int *err;
    int n, T, Tpad;
    int pub;           // steps between two publications
};

template <int D, int MAXW, bool CHAIN>
__global__ __launch_bounds__(64 * MAXW) void k_gsweep_fwd(const double *coef, const double *xin, double *xout,
                                                          GSkew k, double ax, double ay, GFChain ch)
{
    __shared__ double xb[2][MAXW][2];
    int t = blockIdx.x;
    int *fl_in = nullptr, *fl_out = nullptr;  // predecessor's / own progress flag
    if constexpr (CHAIN) {
        const int stage = (int)blockIdx.x / ch.Tpad;
        t = (int)blockIdx.x - stage * ch.Tpad;
        if (t >= ch.T) return;
        xin = stage == 0 ? ch.x0 : ch.xs + (size_t)(stage - 1) * ch.sstride;
        xout = ch.xs + (size_t)stage * ch.sstride;
        if (stage > 0) fl_in = ch.flags + (size_t)((stage - 1) * ch.Tpad + t) * kChainFlagStride;
        if (stage < ch.n - 1) fl_out = ch.flags + (size_t)(stage * ch.Tpad + t) * kChainFlagStride;
    }
    const int r = threadIdx.x, lane = r & 63, wv = r >> 6;
    const int h = k.h, w = k.w, hs = k.hs;
    const bool row_ok = r < h, down_ok = r + 1 < h;
    const unsigned pb = (unsigned)(k.plane * 8);
    const auto rK = make_rsrc(coef + (size_t)t * 6 * k.plane, 6 * pb);
    const auto rI = make_rsrc(xin + (size_t)t * 2 * k.plane, 2 * pb);
    const auto rO = make_rsrc(xout + (size_t)t * 2 * k.plane, 2 * pb);
    static_assert(D >= 2, "a ring slot is refilled while another is in use");
    GFwdSlot S[D];
    auto issue = [&](GFwdSlot &s, int d) {
        const unsigned q = slot_offset(row_ok, d, r, w, hs);
        load_coef6(rK, q, s.K);
        const F64x2 o = load_pair(rI, q);
        s.ou = o.lo;
        s.ov = o.hi;
        // (r + 1, c): lane 63 reads it (the next wave's lane 0 holds it),
        // the other lanes' offset is out of range (no memory access)
        const unsigned qd = lane == 63 ? slot_offset(down_ok, d + 1, r + 1, w, hs) : kOutOfRange;
        const F64x2 dd = load_pair(rI, qd);
        s.du = dd.lo;
        s.dv = dd.hi;
    };
    const int nd = h + w - 1;
    // Chain: the predecessor stores diagonal e at its step e; step d loads
    // diagonals d + D - 1 and (lane 63) d + D, so it needs d + D + 1
    // completed steps (of nd).
    int avail = 0, fpre = 0, pcount = 0;
    if constexpr (CHAIN) {
        pcount = ch.pub;
        if (fl_in) avail = chain_acquire(fl_in, min(D, nd), 0, ch.err);
    }
#pragma unroll
    for (int s = 0; s < D - 1; ++s) issue(S[s], s);
    double pu = 0.0, pv = 0.0;  // this lane's results of the previous step
    for (int d0 = 0; d0 < nd; d0 += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int d = d0 + s;
            if constexpr (CHAIN) {
                if (fl_in) {
                    const int need = min(d + D + 1, nd);
                    if (avail < need) avail = chain_acquire(fl_in, need, __builtin_amdgcn_readfirstlane(fpre), ch.err);
                    fpre = __hip_atomic_load(fl_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            issue(S[(s + D - 1) % D], d + D - 1);
            const GFwdSlot &c = S[s];
            const GFwdSlot &nx = S[(s + 1) % D];  // diagonal d + 1
            const double ru = nx.ou, rv = nx.ov;  // (r, c + 1), old
            const double du_n = wave_from_next_lane(nx.ou), dv_n = wave_from_next_lane(nx.ov);
            const double cdu = lane == 63 ? c.du : du_n, cdv = lane == 63 ? c.dv : dv_n;  // (r + 1, c), old
            double uu = wave_from_prev_lane(pu), uv = wave_from_prev_lane(pv);
            const int wu_src = wv > 0 ? wv - 1 : 0;
            const double eu = xb[(d + 1) & 1][wu_src][0], ev = xb[(d + 1) & 1][wu_src][1];
            if (lane == 0) {
                uu = wv > 0 ? eu : 0.0;
                uv = wv > 0 ? ev : 0.0;
            }
            const unsigned q = slot_offset(row_ok, d, r, w, hs);
            const double A12 = c.K[0], Bu = c.K[1], Bv = c.K[2], wu = c.K[3], wvv = c.K[4], b = c.K[5];
            const double nu = Bu + ax * (pu + ru) + ay * (cdu + uu) - A12 * c.ov;
            double un = wu * nu + (1.0 - kOmega + wu * b) * c.ou;
            const double nv = Bv + ax * (pv + rv) + ay * (cdv + uv) - A12 * un;
            double vn = wvv * nv + (1.0 - kOmega + wvv * b) * c.ov;
            const bool ok = q != kOutOfRange;
            un = ok ? un : 0.0;
            vn = ok ? vn : 0.0;
            store_pair(un, vn, rO, q);
            pu = un;
            pv = vn;
            if (lane == 63) {
                xb[d & 1][wv][0] = un;
                xb[d & 1][wv][1] = vn;
            }
            if constexpr (CHAIN) {
                const bool pub = fl_out && (--pcount == 0 || d + 1 == nd);
                if (pcount == 0) pcount = ch.pub;
                if (pub) wait_stores();
                lds_barrier();
                if (pub && r == 0)
                    __hip_atomic_store(fl_out, min(d + 1, nd), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                lds_barrier();
            }
        }
    }
}

// P consecutive sweeps of k_gsweep_fwd in one wavefront pass (the sweeps
// of a lag block share their coefficients): at step t lane r updates cell
// (r, t - r - 2p) in sweep p.  Sweep p reads sweep p - 1's values of the cell
// (this lane, step t - 2), of its right neighbour (this lane, step t - 1) and
// of its lower neighbour (lane r + 1, step t - 1), and sweep p's values of its
// left (this lane) and upper (lane r - 1) neighbours from step t - 1; sweep 0
// reads the pass input.  Each update is k_gsweep_fwd's arithmetic, so every
// sweep's values equal those of P single-sweep launches.  Coefficients of
// diagonal x sit in ring slot x % NK and the input values of diagonal x in
// slot x % NX, loaded A steps ahead; the step loop is unrolled over NK
// steps (a multiple of NX) so that every slot index is static.
template <int P>
struct GMulti {
    static constexpr int A = 3;                                  // load distance
    static constexpr int NX = A + 1;                             // input ring: diagonals t .. t + A
    static constexpr int NK = (2 * P + A - 1 + NX - 1) / NX * NX;  // coefficients: t - 2(P - 1) .. t + A
    double K[NK][6];
    double X[NX][4];      // ou, ov, and (lane 63) du, dv of the lower neighbour
    double R1[P][2], R2[P][2];  // sweep p's values of steps t - 1 and t - 2
};

template <int P, int MAXW, bool ALL, int S>
__device__ __forceinline__ void gmulti_step(GMulti<P> &st, int t, int r, int lane, int wv, int nw, const GSkew &k,
                                            double ax, double ay, __amdgpu_buffer_rsrc_t rK,
                                            __amdgpu_buffer_rsrc_t rI, const __amdgpu_buffer_rsrc_t *rO, unsigned pb,
                                            double (*xu)[MAXW][P][2], double (*xd)[MAXW][P][2])
{
    using M = GMulti<P>;
    constexpr int NK = M::NK, NX = M::NX, A = M::A;
    const int h = k.h, w = k.w, hs = k.hs;
    const bool row_ok = r < h, down_ok = r + 1 < h;
    {
        // diagonal t + A into the rings
        const int d = t + A;
        const unsigned q = slot_offset(row_ok, d, r, w, hs);
        load_coef6(rK, q, st.K[(S + A) % NK]);
        double *x = st.X[(S + A) % NX];
        const F64x2 o = load_pair(rI, q);
        x[0] = o.lo;
        x[1] = o.hi;
        const unsigned qd = lane == 63 ? slot_offset(down_ok, d + 1, r + 1, w, hs) : kOutOfRange;
        const F64x2 dd = load_pair(rI, qd);
        x[2] = dd.lo;
        x[3] = dd.hi;
    }
    const double *x0 = st.X[S % NX], *x1 = st.X[(S + 1) % NX];  // input of diagonals t, t + 1
    const int pr = (t + 1) & 1;  // LDS half written at step t - 1
    const int wu_src = wv > 0 ? wv - 1 : 0, wd_src = wv + 1 < nw ? wv + 1 : wv;
    double nw_u[P], nw_v[P];
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double *Kp = st.K[((S - 2 * p) % NK + NK) % NK];
        double ou, ov, ru, rv, cdu, cdv;
        if (p == 0) {
            ou = x0[0];
            ov = x0[1];
            ru = x1[0];
            rv = x1[1];
            const double du_n = wave_from_next_lane(x1[0]), dv_n = wave_from_next_lane(x1[1]);
            cdu = lane == 63 ? x0[2] : du_n;
            cdv = lane == 63 ? x0[3] : dv_n;
        } else {
            ou = st.R2[p > 0 ? p - 1 : 0][0];
            ov = st.R2[p > 0 ? p - 1 : 0][1];
            ru = st.R1[p > 0 ? p - 1 : 0][0];
            rv = st.R1[p > 0 ? p - 1 : 0][1];
            cdu = wave_from_next_lane(ru);
            cdv = wave_from_next_lane(rv);
            if (lane == 63) {
                cdu = wv + 1 < nw ? xd[pr][wd_src][p - 1][0] : 0.0;
                cdv = wv + 1 < nw ? xd[pr][wd_src][p - 1][1] : 0.0;
            }
        }
        const double pu = st.R1[p][0], pv = st.R1[p][1];  // (r, c - 1), new
        double uu = wave_from_prev_lane(pu), uv = wave_from_prev_lane(pv);
        if (lane == 0) {
            uu = wv > 0 ? xu[pr][wu_src][p][0] : 0.0;
            uv = wv > 0 ? xu[pr][wu_src][p][1] : 0.0;
        }
        const int d = t - 2 * p;
        const unsigned q = slot_offset(row_ok, d, r, w, hs);
        const double A12 = Kp[0], Bu = Kp[1], Bv = Kp[2], wu = Kp[3], wvv = Kp[4], b = Kp[5];
        const double nu = Bu + ax * (pu + ru) + ay * (cdu + uu) - A12 * ov;
        double un = wu * nu + (1.0 - kOmega + wu * b) * ou;
        const double nv = Bv + ax * (pv + rv) + ay * (cdv + uv) - A12 * un;
        double vn = wvv * nv + (1.0 - kOmega + wvv * b) * ov;
        const bool ok = q != kOutOfRange;
        un = ok ? un : 0.0;
        vn = ok ? vn : 0.0;
        if (ALL || p == P - 1) {
            store_pair(un, vn, rO[ALL ? p : 0], q);
        }
        nw_u[p] = un;
        nw_v[p] = vn;
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
        st.R2[p][0] = st.R1[p][0];
        st.R2[p][1] = st.R1[p][1];
        st.R1[p][0] = nw_u[p];
        st.R1[p][1] = nw_v[p];
    }
    const int pw = t & 1;
    if (lane == 63) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xu[pw][wv][p][0] = nw_u[p];
            xu[pw][wv][p][1] = nw_v[p];
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xd[pw][wv][p][0] = nw_u[p];
            xd[pw][wv][p][1] = nw_v[p];
        }
    }
    lds_barrier();
}

template <int P, int MAXW, bool ALL, int... S>
__device__ __forceinline__ void gmulti_block(GMulti<P> &st, int t0, int r, int lane, int wv, int nw, const GSkew &k,
                                             double ax, double ay, __amdgpu_buffer_rsrc_t rK,
                                             __amdgpu_buffer_rsrc_t rI, const __amdgpu_buffer_rsrc_t *rO, unsigned pb,
                                             double (*xu)[MAXW][P][2], double (*xd)[MAXW][P][2],
                                             std::integer_sequence<int, S...>)
{
    (gmulti_step<P, MAXW, ALL, S>(st, t0 + S, r, lane, wv, nw, k, ax, ay, rK, rI, rO, pb, xu, xd), ...);
}

// ALL: sweep p's values go to xout + p ostride (frame t at t * 2 plane
// within each); else only the last sweep's, to xout.
template <int P, int MAXW, bool ALL>
__global__ __launch_bounds__(64 * MAXW) void k_gsweeps_fwd(const double *coef, const double *xin, double *xout,
                                                           size_t ostride, GSkew k, double ax, double ay)
{
    __shared__ double xu[2][MAXW][P][2];  // lane 63's new values, for the wave below's lane 0
    __shared__ double xd[2][MAXW][P][2];  // lane 0's new values, for the wave above's lane 63
    using M = GMulti<P>;
    const int f = blockIdx.x, r = threadIdx.x, lane = r & 63, wv = __builtin_amdgcn_readfirstlane(r >> 6);
    const int nw = (int)(blockDim.x >> 6);
    const unsigned pb = (unsigned)(k.plane * 8);
    const auto rK = make_rsrc(coef + (size_t)f * 6 * k.plane, 6 * pb);
    const auto rI = make_rsrc(xin + (size_t)f * 2 * k.plane, 2 * pb);
    __amdgpu_buffer_rsrc_t rO[ALL ? P : 1];
#pragma unroll
    for (int p = 0; p < (ALL ? P : 1); ++p) rO[p] = make_rsrc(xout + p * ostride + (size_t)f * 2 * k.plane, 2 * pb);
    for (int i = threadIdx.x; i < 2 * MAXW * P * 2; i += blockDim.x) {
        (&xu[0][0][0][0])[i] = 0.0;
        (&xd[0][0][0][0])[i] = 0.0;
    }
    M st;
#pragma unroll
    for (int p = 0; p < P; ++p) st.R1[p][0] = st.R1[p][1] = st.R2[p][0] = st.R2[p][1] = 0.0;
    const bool row_ok = r < k.h, down_ok = r + 1 < k.h;
    // prologue: diagonals 0 .. A - 1
#pragma unroll
    for (int d = 0; d < M::A; ++d) {
        const unsigned q = slot_offset(row_ok, d, r, k.w, k.hs);
        load_coef6(rK, q, st.K[d]);
        const F64x2 o = load_pair(rI, q);
        st.X[d][0] = o.lo;
        st.X[d][1] = o.hi;
        const unsigned qd = lane == 63 ? slot_offset(down_ok, d + 1, r + 1, k.w, k.hs) : kOutOfRange;
        const F64x2 dd = load_pair(rI, qd);
        st.X[d][2] = dd.lo;
        st.X[d][3] = dd.hi;
    }
    __syncthreads();
    const int Tend = (k.h + k.w - 2) + 2 * (P - 1);
    for (int t0 = 0; t0 <= Tend; t0 += M::NK)
        gmulti_block<P, MAXW, ALL>(st, t0, r, lane, wv, nw, k, ax, ay, rK, rI, rO, pb, xu, xd,
                              std::make_integer_sequence<int, M::NK>{});
}

// Reverse-mode derivative of k_gsweep_fwd.  lam_out: adjoint of xout.
// Adds the coefficient adjoints to gK (or stores them, gK_set) and the
// partials of ax, ay to gaxy (T, 2, plane; one sum per row, element r of
// each plane), and writes lam_in, the adjoint of xin.  Diagonals run from
// last to first: a cell's new values are read only by its right and lower
// neighbours (diagonal d + 1) and by the next sweep (lam_out), so the right
// neighbour's adjoint is the lane's own previous result and the lower one
// lane r + 1's.  A cell's old values are read by its own update and by the
// updates of its left and upper neighbours (diagonal d - 1, the next step),
// so its lam_in is complete once the lane and lane r - 1 have passed the
// next diagonal: it is formed two steps later (the upper neighbour's
// adjoint of the wave's lane 0 comes through LDS after a barrier).
// Ring slot of one diagonal (diagonals run last to first).  The new values
// of the left and upper neighbours are the lane's own / lane r - 1's slot
// of the next diagonal processed (d - 1), the old values of the right and
// lower neighbours the lane's own / lane r + 1's old values of the previous
// step (d + 1), carried in registers; only the wave's edge lanes load them
// (uu, uv by lane 0, du, dv by lane 63), so each value is read once.
struct GBwdSlot {
    double K[6], lu_out, lv_out, un, vn, uu, uv, ou, ov, du, dv;
};

// What a cell's lam_in needs of its own step: the adjoints of its update
// (anu, anv, adu, adv) and its coefficients wu, wv, b, A12.
struct GBwdHist {
    double anu, anv, adu, adv, wu, wv, b, A12;
};

// The adjoint sweeps of one lag block as a chain (CHAIN): work-group
// (stage s, frame t) runs the adjoint of the block's sweep n - 1 - s, the
// same code as a single launch, with its lam_out being stage s - 1's lam_in
// and its coefficient adjoints added to what stage s - 1 stored.  Stage s
// waits, before it loads the lam_out of diagonal x, for the step after which
// that diagonal's lam_in is complete.  The row partials of ax, ay are added
// stage after stage (a last flag value), so every sum is formed in the order
// of single launches: the chain's results are bit-identical to n single
// adjoint launches.
struct GChain {
    const double *x0;       // the block's start state, x_in of its first sweep
    const double *xs;       // state after the block's sweep j at xs + j sstride
    size_t sstride;
    const double *lam_top;  // adjoint of the block's last state
    double *lam_mid;        // lam_in of stage s < n - 1 at lam_mid + s lstride
    size_t lstride;
    double *lam_bottom;     // adjoint of the block's start state
    int *flags;             // stage s, frame t: flags[(s Tpad + t) kChainFlagStride]
    int *err;
    int n, T, Tpad;
    int pub;                // steps between two publications
};

template <int D, int MAXW, bool CHAIN>
__global__ __launch_bounds__(64 * MAXW) void k_gsweep_bwd(const double *coef, const double *xin,
                                                          const double *xout, const double *lam_out,
                                                          double *gK, double *gaxy, double *lam_in, GSkew k,
                                                          double ax, double ay, int gK_set, GChain ch)
{
    __shared__ double xb[2][MAXW][2];  // lane 0's anu, anv, for the wave above's lane 63
    __shared__ double xu[2][MAXW][2];  // lane 63's anu, anv, for the wave below's lane 0
    int t = blockIdx.x;
    int *fl_in = nullptr, *fl_out = nullptr;  // producer's / own progress flag
    if constexpr (CHAIN) {
        const int stage = (int)blockIdx.x / ch.Tpad, sw = ch.n - 1 - stage;
        t = (int)blockIdx.x - stage * ch.Tpad;
        if (t >= ch.T) return;
        xin = sw == 0 ? ch.x0 : ch.xs + (size_t)(sw - 1) * ch.sstride;
        xout = ch.xs + (size_t)sw * ch.sstride;
        lam_out = stage == 0 ? ch.lam_top : ch.lam_mid + (size_t)(stage - 1) * ch.lstride;
        lam_in = stage == ch.n - 1 ? ch.lam_bottom : ch.lam_mid + (size_t)stage * ch.lstride;
        gK_set = gK_set && stage == 0;
        if (stage > 0) fl_in = ch.flags + (size_t)((stage - 1) * ch.Tpad + t) * kChainFlagStride;
        if (stage < ch.n - 1) fl_out = ch.flags + (size_t)(stage * ch.Tpad + t) * kChainFlagStride;
    }
    const int r = threadIdx.x, lane = r & 63, wv = r >> 6, nw = (int)(blockDim.x >> 6);
    const int h = k.h, w = k.w, hs = k.hs;
    const bool row_ok = r < h, down_ok = r + 1 < h, up_ok = r >= 1 && r - 1 < h;
    const unsigned pb = (unsigned)(k.plane * 8);
    const auto rK = make_rsrc(coef + (size_t)t * 6 * k.plane, 6 * pb);
    const auto rI = make_rsrc(xin + (size_t)t * 2 * k.plane, 2 * pb);
    const auto rO = make_rsrc(xout + (size_t)t * 2 * k.plane, 2 * pb);
    const auto rL = make_rsrc(lam_out + (size_t)t * 2 * k.plane, 2 * pb);
    const auto rG = make_rsrc(gK + (size_t)t * 6 * k.plane, 6 * pb);
    const auto rN = make_rsrc(lam_in + (size_t)t * 2 * k.plane, 2 * pb);
    static_assert(D >= 2 && D % 2 == 0, "the gK ring alternates with the step parity");
    GBwdSlot S[D];
    double Gc[2][6];  // gK of the next step's cell, loaded one step ahead
    GBwdHist Hs[2];   // the cells of the last two steps, by step parity
    Hs[0] = Hs[1] = GBwdHist{0, 0, 0, 0, 0, 0, 0, 0};
    auto issue = [&](GBwdSlot &s, int d) {
        const unsigned q = slot_offset(row_ok, d, r, w, hs);
        load_coef6(rK, q, s.K);
        const F64x2 l = load_pair(rL, q), n = load_pair(rO, q), o = load_pair(rI, q);
        s.lu_out = l.lo;
        s.lv_out = l.hi;
        s.un = n.lo;
        s.vn = n.hi;
        s.ou = o.lo;
        s.ov = o.hi;
        const unsigned qu = lane == 0 ? slot_offset(up_ok, d - 1, r - 1, w, hs) : kOutOfRange;     // (r - 1, c), new
        const F64x2 uu = load_pair(rO, qu);
        s.uu = uu.lo;
        s.uv = uu.hi;
        const unsigned qd = lane == 63 ? slot_offset(down_ok, d + 1, r + 1, w, hs) : kOutOfRange;  // (r + 1, c), old
        const F64x2 dd = load_pair(rI, qd);
        s.du = dd.lo;
        s.dv = dd.hi;
    };
    auto issue_g = [&](double *g, int d) {
        // gK_set: the sweep stores its coefficient adjoints instead of
        // adding them (out-of-range loads read 0, without a memory access)
        const unsigned q = gK_set ? kOutOfRange : slot_offset(row_ok, d, r, w, hs);
        load_coef6(rG, q, g);
    };
    const int nd = h + w - 1;
    // Chain: after its step i the producer has stored the lam_in of
    // diagonals >= nd + 1 - i and the coefficient adjoints of diagonals
    // >= nd - 1 - i; step i loads lam_out of diagonal nd - i - D, so it
    // needs i + D + 2 completed steps (of nd + 2 in all).
    int avail = 0, fpre = 0, pcount = 0;
    if constexpr (CHAIN) {
        pcount = ch.pub;
        if (fl_in) avail = chain_acquire(fl_in, min(D + 1, nd + 2), 0, ch.err);
    }
    // step i handles diagonal nd - 1 - i; steps nd and nd + 1 only finish
    // the lam_in of the last two diagonals
#pragma unroll
    for (int s = 0; s < D - 1; ++s) issue(S[s], nd - 1 - s);
    issue_g(Gc[0], nd - 1);
    double pnu = 0.0, pnv = 0.0;  // this lane's anu, anv of the previous step
    double pou = 0.0, pov = 0.0;  // this lane's old values of the previous step's diagonal (d + 1)
    double g_axr = 0.0, g_ayr = 0.0;  // this row's partials of ax, ay
    for (int i0 = 0; i0 < nd + 2; i0 += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int i = i0 + s, d = nd - 1 - i;
            // lam_in of the cell of step i - 2 (diagonal d + 2, column c2):
            // its left neighbour's anu is this lane's of step i - 1 (pnu),
            // its upper neighbour's lane r - 1's of step i - 1
            {
                const GBwdHist &H = Hs[s & 1];  // step i - 2 (D is even)
                double unu = wave_from_prev_lane(pnu), unv = wave_from_prev_lane(pnv);
                const int wu_src = wv > 0 ? wv - 1 : 0;
                const double eu = xu[(i + 1) & 1][wu_src][0], ev = xu[(i + 1) & 1][wu_src][1];
                if (lane == 0) {
                    unu = wv > 0 ? eu : 0.0;
                    unv = wv > 0 ? ev : 0.0;
                }
                const int c2 = d + 2 - r;
                const unsigned q2 = slot_offset(row_ok, d + 2, r, w, hs);
                const double Lu = (1.0 - kOmega + H.wu * H.b) * H.adu + (c2 > 0 ? ax * pnu : 0.0) +
                                  (r > 0 ? ay * unu : 0.0);
                const double Lv = (1.0 - kOmega + H.wv * H.b) * H.adv + (c2 > 0 ? ax * pnv : 0.0) +
                                  (r > 0 ? ay * unv : 0.0) - H.A12 * H.anu;
                store_pair(Lu, Lv, rN, q2);
            }
            if constexpr (CHAIN) {
                if (fl_in) {
                    const int need = min(i + D + 2, nd + 2);
                    if (avail < need) avail = chain_acquire(fl_in, need, __builtin_amdgcn_readfirstlane(fpre), ch.err);
                    fpre = __hip_atomic_load(fl_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
            issue_g(Gc[(s + 1) & 1], d - 1);  // D is even: the parity of s is the parity of i
            issue(S[(s + D - 1) % D], d - (D - 1));
            const GBwdSlot &c = S[s];
            const GBwdSlot &nx = S[(s + 1) % D];  // diagonal d - 1
            const double lu = nx.un, lv = nx.vn;  // (r, c - 1), new
            const double uu_p = wave_from_prev_lane(nx.un), uv_p = wave_from_prev_lane(nx.vn);
            const double cuu = lane == 0 ? c.uu : uu_p, cuv = lane == 0 ? c.uv : uv_p;  // (r - 1, c), new
            const double ru = pou, rv = pov;  // (r, c + 1), old
            const double du_n = wave_from_next_lane(pou), dv_n = wave_from_next_lane(pov);
            const double cdu = lane == 63 ? c.du : du_n, cdv = lane == 63 ? c.dv : dv_n;  // (r + 1, c), old
            pou = c.ou;
            pov = c.ov;
            double dnu = wave_from_next_lane(pnu), dnv = wave_from_next_lane(pnv);
            const int wd_src = wv + 1 < nw ? wv + 1 : wv;
            const double eu = xb[(i + 1) & 1][wd_src][0], ev = xb[(i + 1) & 1][wd_src][1];
            if (lane == 63) {
                dnu = wv + 1 < nw ? eu : 0.0;
                dnv = wv + 1 < nw ? ev : 0.0;
            }
            const unsigned q = slot_offset(row_ok, d, r, w, hs);
            const bool ok = q != kOutOfRange;
            const double A12 = c.K[0], Bu = c.K[1], Bv = c.K[2], wu = c.K[3], wvv = c.K[4], b = c.K[5];
            double a_dv = c.lv_out + ax * pnv + ay * dnv;
            double a_nv = wvv * a_dv;
            double a_du = c.lu_out + ax * pnu + ay * dnu - A12 * a_nv;
            double a_nu = wu * a_du;
            a_dv = ok ? a_dv : 0.0;
            a_nv = ok ? a_nv : 0.0;
            a_du = ok ? a_du : 0.0;
            a_nu = ok ? a_nu : 0.0;
            const double nu = Bu + ax * (lu + ru) + ay * (cdu + cuu) - A12 * c.ov;
            const double nv = Bv + ax * (lv + rv) + ay * (cdv + cuv) - A12 * c.un;
            const double *g = Gc[s & 1];
            const double gn[6] = {g[0] - c.ov * a_nu - c.un * a_nv,          // A12
                                  g[1] + a_nu,                               // Bu
                                  g[2] + a_nv,                               // Bv
                                  g[3] + (nu + b * c.ou) * a_du,             // wu
                                  g[4] + (nv + b * c.ov) * a_dv,             // wv
                                  g[5] + wu * c.ou * a_du + wvv * c.ov * a_dv};  // b
            store_coef6(gn, rG, q);
            g_axr += (lu + ru) * a_nu + (lv + rv) * a_nv;
            g_ayr += (cdu + cuu) * a_nu + (cdv + cuv) * a_nv;
            Hs[s & 1] = GBwdHist{a_nu, a_nv, a_du, a_dv, wu, wvv, b, A12};  // read again at step i + 2
            pnu = a_nu;
            pnv = a_nv;
            if (lane == 0) {
                xb[i & 1][wv][0] = a_nu;
                xb[i & 1][wv][1] = a_nv;
            }
            if (lane == 63) {
                xu[i & 1][wv][0] = a_nu;
                xu[i & 1][wv][1] = a_nv;
            }
            if constexpr (CHAIN) {
                const bool pub = fl_out && (--pcount == 0 || i + 1 == nd + 2);
                if (pcount == 0) pcount = ch.pub;
                if (pub) wait_stores();
                lds_barrier();
                if (pub && r == 0)
                    __hip_atomic_store(fl_out, min(i + 1, nd + 2), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                lds_barrier();
            }
        }
    }
    if constexpr (CHAIN) {
        if (fl_in) chain_acquire(fl_in, nd + 3, __builtin_amdgcn_readfirstlane(fpre), ch.err);  // producer's row sums added
    }
    if (row_ok) {
        double *G = gaxy + (size_t)t * 2 * k.plane;
        G[r] += g_axr;
        G[k.plane + r] += g_ayr;
    }
    if constexpr (CHAIN) {
        if (fl_out) {
            wait_stores();
            __syncthreads();
            if (r == 0) __hip_atomic_store(fl_out, nd + 3, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// --------------------------------------------------------------- median
// scipy median_filter(size=(5, 5), mode="mirror") of du and dv (interior,
// core/optical_flow.py:607-609) and its derivative: the gradient of each
// output goes to the sample that the selection returned.
__device__ inline int mirror5(int j, int n)
{
    if (n == 1) return 0;
    const int period = 2 * n - 2;
    j %= period;
    if (j < 0) j += period;
    return j < n ? j : period - j;
}

__global__ __launch_bounds__(64) void k_gmedian(const double *x, int T, int h, int w, double *out,
                                                 const double *gout, double *gx)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, tq = blockIdx.z;
    if (c >= w) return;
    const size_t hw = (size_t)h * w;
    const double *X = x + (size_t)tq * hw;
    int rr[5], cc[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        rr[k] = mirror5(r + k - 2, h);
        cc[k] = mirror5(c + k - 2, w);
    }
    double v[25], s[25];
#pragma unroll
    for (int i = 0; i < 5; ++i)
#pragma unroll
        for (int j = 0; j < 5; ++j) v[i * 5 + j] = s[i * 5 + j] = X[(size_t)rr[i] * w + cc[j]];
    const double m = median25(s);
    if (out) out[(size_t)tq * hw + (size_t)r * w + c] = m;
    if (gx) {
        // the first window sample equal to the median receives the gradient
        int ki = 0, kj = 0;
#pragma unroll
        for (int k = 24; k >= 0; --k)
            if (v[k] == m) {
                ki = k / 5;
                kj = k % 5;
            }
        int sr = rr[0], sc = cc[0];
#pragma unroll
        for (int k = 1; k < 5; ++k) {
            sr = ki == k ? rr[k] : sr;
            sc = kj == k ? cc[k] : sc;
        }
        atomicAdd(gx + (size_t)tq * hw + (size_t)sr * w + sc, gout[(size_t)tq * hw + (size_t)r * w + c]);
    }
}

// ------------------------------------------------------ continuous warp
// core/warping.py:160-234 with a continuous map: map = x + u, the
// out-of-bounds test on the unclipped map, clip to [0, W-1], Keys cubic
// (A = -0.75) with replicated borders; out-of-bounds pixels take f1.
__device__ inline void keys(double t, double *wt, double *dwt)
{
    const double A = -0.75;
    const double x[4] = {1.0 + t, t, 1.0 - t, 2.0 - t};
    const double s[4] = {1.0, 1.0, -1.0, -1.0};  // d|x|/dt
    for (int i = 0; i < 4; ++i) {
        const double a = x[i];
        if (a <= 1.0) {
            wt[i] = ((A + 2) * a - (A + 3)) * a * a + 1;
            dwt[i] = s[i] * ((3 * (A + 2) * a - 2 * (A + 3)) * a);
        } else {
            wt[i] = ((A * a - 5 * A) * a + 8 * A) * a - 4 * A;
            dwt[i] = s[i] * ((3 * A * a - 10 * A) * a + 8 * A);
        }
    }
}

__global__ __launch_bounds__(64) void k_gwarp(const double *img, const double *fill, int fill_per_frame, const double *u,
                        const double *v, int T, int C, int H, int W, double *out, const double *gout,
                        double *gimg, double *gu, double *gv)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, t = blockIdx.z;
    if (x >= W) return;
    const size_t HW = (size_t)H * W, p = (size_t)y * W + x;
    const double mx = x + u[(size_t)t * HW + p], my = y + v[(size_t)t * HW + p];
    const bool oob = mx < 0 || mx >= W || my < 0 || my >= H;
    const double *I = img + (size_t)t * C * HW;
    if (oob) {
        if (out)
            for (int c = 0; c < C; ++c)
                out[((size_t)t * C + c) * HW + p] = fill[(fill_per_frame ? (size_t)t * C : 0) * HW + c * HW + p];
        if (gout && gu) {  // no dependence on img, u or v
            gu[(size_t)t * HW + p] = 0.0;
            gv[(size_t)t * HW + p] = 0.0;
        }
        return;
    }
    const bool clipx = mx > W - 1, clipy = my > H - 1;
    const double cx = clipx ? W - 1.0 : mx, cy = clipy ? H - 1.0 : my;
    const double fx0 = floor(cx), fy0 = floor(cy);
    double wx[4], dwx[4], wy[4], dwy[4];
    keys(cx - fx0, wx, dwx);
    keys(cy - fy0, wy, dwy);
    int xs[4], ys[4];
    for (int i = 0; i < 4; ++i) {
        int xx = (int)fx0 - 1 + i, yy = (int)fy0 - 1 + i;
        xs[i] = xx < 0 ? 0 : (xx > W - 1 ? W - 1 : xx);
        ys[i] = yy < 0 ? 0 : (yy > H - 1 ? H - 1 : yy);
    }
    double gsu = 0.0, gsv = 0.0;
    for (int c = 0; c < C; ++c) {
        const double *S = I + c * HW;
        if (out) {
            double s = 0.0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) s += wy[i] * wx[j] * S[(size_t)ys[i] * W + xs[j]];
            out[((size_t)t * C + c) * HW + p] = s;
        }
        if (gout) {
            const double g = gout[((size_t)t * C + c) * HW + p];
            if (g == 0.0) continue;
            double sx = 0.0, sy = 0.0;
            for (int i = 0; i < 4; ++i)
                for (int j = 0; j < 4; ++j) {
                    const double val = S[(size_t)ys[i] * W + xs[j]];
                    sx += wy[i] * dwx[j] * val;
                    sy += dwy[i] * wx[j] * val;
                    if (gimg) atomicAdd(gimg + ((size_t)t * C + c) * HW + (size_t)ys[i] * W + xs[j],
                                        g * wy[i] * wx[j]);
                }
            gsu += g * sx;
            gsv += g * sy;
        }
    }
    if (gout && gu) {
        gu[(size_t)t * HW + p] = clipx ? 0.0 : gsu;
        gv[(size_t)t * HW + p] = clipy ? 0.0 : gsv;
    }
}

}  // namespace grad
}  // namespace frh

using namespace frh;
using namespace frh::grad;

// Chains: at most kChainMax sweeps per launch (longer blocks take several
// launches).  Scratch of the chained adjoint: the failure word, the flags,
// then the lam_in of every stage but the last; of the chained sweeps: the
// failure word and the flags.
constexpr int kChainMax = 32;


static size_t chain_scratch_bytes(int n, int T, const GSkew &k)
{
    const int nc = n < kChainMax ? n : kChainMax;
    const int planes = nc - 1 + (n > kChainMax ? 1 : 0);  // + the hand-over between launches
    return 256 + (size_t)nc * chain_tpad(T) * kChainFlagStride * 4 + (size_t)planes * T * 2 * k.plane * 8;
}


// Sweeps per pipelined launch; a lag block of more sweeps takes several.
// At 512 lanes (two waves per SIMD, 256 registers) 5 sweeps spill
// registers and run slower than 3 + 2.
__host__ __device__ constexpr int gmax_pipe(int maxw) { return maxw <= 4 ? 5 : 4; }

template <int MAXW, bool ALL>
static void launch_gsweeps(int P, const double *coef, const double *x_in, double *x_out, size_t ostride,
                           int T, const GSkew &k, double ax, double ay, hipStream_t st)
{
    const dim3 g(T), b(k.hs);
    switch (P) {
    case 2:
        hipLaunchKernelGGL((k_gsweeps_fwd<2, MAXW, ALL>), g, b, 0, st, coef, x_in, x_out, ostride, k, ax, ay);
        break;
    case 3:
        hipLaunchKernelGGL((k_gsweeps_fwd<3, MAXW, ALL>), g, b, 0, st, coef, x_in, x_out, ostride, k, ax, ay);
        break;
    case 4:
        hipLaunchKernelGGL((k_gsweeps_fwd<4, MAXW, ALL>), g, b, 0, st, coef, x_in, x_out, ostride, k, ax, ay);
        break;
    default:
        if constexpr (gmax_pipe(MAXW) >= 5)
            hipLaunchKernelGGL((k_gsweeps_fwd<5, MAXW, ALL>), g, b, 0, st, coef, x_in, x_out, ostride, k, ax, ay);
        break;
    }
}

static void launch_gsweep1(const double *coef, const double *x_in, double *x_out, int T, const GSkew &k,
                           double ax, double ay, hipStream_t st)
{
    const dim3 b(k.hs);
    if (k.hs <= 256)
        hipLaunchKernelGGL((k_gsweep_fwd<4, 4, false>), dim3(T), b, 0, st, coef, x_in, x_out, k, ax, ay, GFChain{});
    else if (k.hs <= 512)
        hipLaunchKernelGGL((k_gsweep_fwd<4, 8, false>), dim3(T), b, 0, st, coef, x_in, x_out, k, ax, ay, GFChain{});
    else
        hipLaunchKernelGGL((k_gsweep_fwd<2, 16, false>), dim3(T), b, 0, st, coef, x_in, x_out, k, ax, ay, GFChain{});
}

// A lag block's sweeps as one chain launch (at most kChainMax sweeps).
static void launch_gsweep_chain(const double *coef, const GFChain &ch, const GSkew &k, double ax, double ay,
                                hipStream_t st)
{
    const dim3 g(ch.n * ch.Tpad), b(k.hs);
    if (k.hs <= 256)
        hipLaunchKernelGGL((k_gsweep_fwd<4, 4, true>), g, b, 0, st, coef, nullptr, nullptr, k, ax, ay, ch);
    else if (k.hs <= 512)
        hipLaunchKernelGGL((k_gsweep_fwd<4, 8, true>), g, b, 0, st, coef, nullptr, nullptr, k, ax, ay, ch);
    else
        hipLaunchKernelGGL((k_gsweep_fwd<2, 16, true>), g, b, 0, st, coef, nullptr, nullptr, k, ax, ay, ch);
}

extern "C" {

static int grad_fail(const char *what, hipError_t e)
{
    return frh_report_error(FRH_EHIP, "%s: %s", what, hipGetErrorString(e));
}

#define FRH_GCHECK(cond, msg)                                   \
    do {                                                        \
        if (!(cond)) return frh_report_error(FRH_EINVAL, "%s", msg); \
    } while (0)
#define FRH_GLAUNCH(what)                                       \
    do {                                                        \
        hipError_t e_ = hipGetLastError();                      \
        if (e_ != hipSuccess) return grad_fail(what, e_);       \
    } while (0)

int frh_grad_tensor(const double *fixed, const double *moving, int T, int C, int h, int w, double hy,
                    double hx, double *J, const double *gJ, double *g_moving, void *stream)
{
    FRH_GCHECK(fixed && moving && T > 0 && C > 0 && C <= kMaxC && h > 0 && w > 0 && T * C <= 65535 &&
                   h <= 65535, "frh_grad_tensor: bad arguments (at most 2 channels)");
    hipStream_t st = (hipStream_t)stream;
    dim3 g((w + 63) / 64, h, T * C), b(64);
    if (J) {
        hipLaunchKernelGGL(k_gtensor_fwd, g, b, 0, st, fixed, moving, T, C, h, w, hy, hx, J);
        FRH_GLAUNCH("tensor forward");
    }
    if (gJ) {
        FRH_GCHECK(g_moving != nullptr, "frh_grad_tensor: g_moving is null");
        hipLaunchKernelGGL(k_gtensor_bwd, g, b, 0, st, fixed, moving, gJ, T, C, h, w, hy, hx, g_moving);
        FRH_GLAUNCH("tensor backward");
    }
    return FRH_OK;
}

static bool make_gskew(int h, int w, GSkew &k)
{
    if (h <= 0 || w <= 0 || h > 1024 || w > 65535) return false;
    k.h = h;
    k.w = w;
    k.hs = 64 * ((h + 63) / 64);
    k.plane = (size_t)(h + w - 1) * (size_t)k.hs;
    return 6 * k.plane * 8 < (size_t)0x7fffffff;  // one frame's coefficients within a buffer resource
}

int frh_grad_skew_plane(int h, int w, size_t *plane)
{
    GSkew k;
    FRH_GCHECK(plane && make_gskew(h, w, k), "frh_grad_skew_plane: bad level size (h <= 1024)");
    *plane = k.plane;
    return FRH_OK;
}

int frh_grad_skew(const double *src, int N, int K, int h, int w, double *dst, int to_skew, void *stream)
{
    GSkew k;
    FRH_GCHECK(src && dst && N > 0 && N <= 65535 && K > 0 && K <= 64 && make_gskew(h, w, k),
               "frh_grad_skew: bad arguments");
    hipLaunchKernelGGL(k_gskew, dim3((unsigned)((k.plane + 255) / 256), N), dim3(256), 0, (hipStream_t)stream,
                       src, dst, K, k, to_skew);
    FRH_GLAUNCH("skew conversion");
    return FRH_OK;
}

int frh_grad_coef(const double *x, const double *u0, const double *J, const double *W, int T, int C,
                  int h, int w, double ax, double ay, const double *a_data, int psi_one, double *coef,
                  const double *g_coef, double *g_x, double *g_u0, double *g_J, double *g_par,
                  void *stream)
{
    GSkew k;
    FRH_GCHECK(x && J && W && a_data && T > 0 && C > 0 && C <= kMaxC && T <= 65535 && make_gskew(h, w, k),
               "frh_grad_coef: bad arguments (at most 2 channels and 1024 rows)");
    FRH_GCHECK(!g_coef || (g_x && g_J && g_par), "frh_grad_coef: g_x, g_J and g_par are required with g_coef");
    hipStream_t st = (hipStream_t)stream;
    dim3 g(coef_tiles(h, w), T), b(kCoefTR * kCoefTC);
    if (coef) {
        if (C == 1)
            hipLaunchKernelGGL(k_gcoef_fwd<1>, g, b, 0, st, x, u0, J, W, k, ax, ay, a_data, psi_one, coef);
        else
            hipLaunchKernelGGL(k_gcoef_fwd<2>, g, b, 0, st, x, u0, J, W, k, ax, ay, a_data, psi_one, coef);
        FRH_GLAUNCH("coefficient forward");
    }
    if (g_coef) {
        if (C == 1)
            hipLaunchKernelGGL(k_gcoef_bwd<1>, g, b, 0, st, x, u0, J, W, g_coef, k, ax, ay, a_data, psi_one, g_x,
                               g_u0, g_J, g_par);
        else
            hipLaunchKernelGGL(k_gcoef_bwd<2>, g, b, 0, st, x, u0, J, W, g_coef, k, ax, ay, a_data, psi_one, g_x,
                               g_u0, g_J, g_par);
        FRH_GLAUNCH("coefficient backward");
    }
    return FRH_OK;
}

int frh_grad_sweep(const double *coef, const double *x_in, double *x_out, int T, int h, int w,
                   double ax, double ay, void *stream)
{
    GSkew k;
    FRH_GCHECK(coef && x_in && x_out && T > 0 && T <= 65535 && make_gskew(h, w, k),
               "frh_grad_sweep: bad arguments (at most 1024 rows)");
    launch_gsweep1(coef, x_in, x_out, T, k, ax, ay, (hipStream_t)stream);
    FRH_GLAUNCH("sweep forward");
    return FRH_OK;
}

size_t frh_grad_sweeps_scratch_bytes(int n_sweeps, int T)
{
    if (n_sweeps <= 0 || T <= 0) return 0;
    const int nc = n_sweeps < kChainMax ? n_sweeps : kChainMax;
    return 256 + (size_t)nc * chain_tpad(T) * kChainFlagStride * 4;
}

int frh_grad_sweeps(const double *coef, const double *x_in, double *x_out, size_t out_stride, int n_sweeps,
                    int write_all, int T, int h, int w, double ax, double ay, void *scratch, size_t scratch_bytes,
                    void *stream)
{
    GSkew k;
    FRH_GCHECK(coef && x_in && x_out && n_sweeps > 0 && T > 0 && T <= 65535 && make_gskew(h, w, k),
               "frh_grad_sweeps: bad arguments (at most 1024 rows)");
    const size_t state = (size_t)T * 2 * k.plane;
    FRH_GCHECK(!write_all || n_sweeps == 1 || out_stride >= state,
               "frh_grad_sweeps: out_stride is smaller than one state (T * 2 * plane values)");
    FRH_GCHECK(!write_all || x_in + state <= x_out || x_in >= x_out + (size_t)(n_sweeps - 1) * out_stride + state,
               "frh_grad_sweeps: x_in must not overlap the written states");
    hipStream_t st = (hipStream_t)stream;
    if (scratch && n_sweeps > 1) {
        // one chain launch per at most kChainMax sweeps
        const size_t need = frh_grad_sweeps_scratch_bytes(n_sweeps, T);
        if (scratch_bytes < need)
            return frh_report_error(FRH_EINVAL, "frh_grad_sweeps: scratch of %zu bytes needed, %zu given", need,
                                    scratch_bytes);
        const int Tpad = (int)chain_tpad(T);
        const double *src = x_in;
        for (int lo = 0; lo < n_sweeps; lo += kChainMax) {
            const int n = n_sweeps - lo < kChainMax ? n_sweeps - lo : kChainMax;
            GFChain ch{};
            ch.x0 = src;
            ch.xs = write_all ? x_out + (size_t)lo * out_stride : x_out;
            ch.sstride = write_all ? out_stride : 0;
            ch.err = (int *)scratch;
            ch.flags = (int *)((char *)scratch + 256);
            ch.n = n;
            ch.T = T;
            ch.Tpad = Tpad;
            ch.pub = chain_publish_interval(k.h + k.w - 1);
            if (hipError_t e = launch_clear_flags(ch.flags, n * Tpad, st); e != hipSuccess)
                return grad_fail("chained sweeps flags", e);
            launch_gsweep_chain(coef, ch, k, ax, ay, st);
            FRH_GLAUNCH("chained sweeps");
            src = ch.xs + (size_t)(n - 1) * ch.sstride;
        }
        return FRH_OK;
    }
    const double *src = x_in;
    for (int j = 0; j < n_sweeps;) {
        int P = n_sweeps - j;
        const int pmax = k.hs > 512 ? 1 : gmax_pipe(k.hs <= 256 ? 4 : 8);  // 1024 lanes: single sweeps
        if (P > pmax) {  // even chunks: 9 sweeps as 5 + 4
            const int nch = (P + pmax - 1) / pmax;
            P = (P + nch - 1) / nch;
        }
        double *dst = write_all ? x_out + (size_t)j * out_stride : x_out;
        if (P == 1)
            launch_gsweep1(coef, src, dst, T, k, ax, ay, st);
        else if (k.hs <= 256)
            write_all ? launch_gsweeps<4, true>(P, coef, src, dst, out_stride, T, k, ax, ay, st)
                      : launch_gsweeps<4, false>(P, coef, src, dst, 0, T, k, ax, ay, st);
        else
            write_all ? launch_gsweeps<8, true>(P, coef, src, dst, out_stride, T, k, ax, ay, st)
                      : launch_gsweeps<8, false>(P, coef, src, dst, 0, T, k, ax, ay, st);
        FRH_GLAUNCH("sweeps forward");
        src = write_all ? dst + (size_t)(P - 1) * out_stride : dst;
        j += P;
    }
    return FRH_OK;
}

int frh_grad_sweep_adjoint(const double *coef, const double *x_in, const double *x_out,
                           const double *lam_out, double *lam_in, double *g_coef, int g_coef_set, double *g_axy,
                           int T, int h, int w, double ax, double ay, void *stream)
{
    GSkew k;
    FRH_GCHECK(coef && x_in && x_out && lam_out && lam_in && g_coef && g_axy && T > 0 && T <= 65535 &&
                   make_gskew(h, w, k),
               "frh_grad_sweep_adjoint: bad arguments (at most 1024 rows)");
    FRH_GCHECK(lam_in != lam_out, "frh_grad_sweep_adjoint: lam_in must not alias lam_out");
    hipStream_t st = (hipStream_t)stream;
    const dim3 b(k.hs);
    const int set = g_coef_set ? 1 : 0;
    if (k.hs <= 256)
        hipLaunchKernelGGL((k_gsweep_bwd<4, 4, false>), dim3(T), b, 0, st, coef, x_in, x_out, lam_out, g_coef, g_axy,
                           lam_in, k, ax, ay, set, GChain{});
    else if (k.hs <= 512)
        hipLaunchKernelGGL((k_gsweep_bwd<4, 8, false>), dim3(T), b, 0, st, coef, x_in, x_out, lam_out, g_coef, g_axy,
                           lam_in, k, ax, ay, set, GChain{});
    else
        hipLaunchKernelGGL((k_gsweep_bwd<2, 16, false>), dim3(T), b, 0, st, coef, x_in, x_out, lam_out, g_coef, g_axy,
                           lam_in, k, ax, ay, set, GChain{});
    FRH_GLAUNCH("sweep adjoint");
    return FRH_OK;
}

size_t frh_grad_sweeps_adjoint_scratch_bytes(int n_sweeps, int T, int h, int w)
{
    GSkew k;
    if (n_sweeps <= 0 || T <= 0 || !make_gskew(h, w, k)) return 0;
    return chain_scratch_bytes(n_sweeps, T, k);
}

int frh_grad_sweeps_adjoint(const double *coef, const double *x0, const double *xs, size_t state_stride,
                            int n_sweeps, const double *lam_top, double *lam_bottom, double *g_coef, int g_coef_set,
                            double *g_axy, int T, int h, int w, double ax, double ay, void *scratch,
                            size_t scratch_bytes, void *stream)
{
    GSkew k;
    FRH_GCHECK(coef && x0 && xs && lam_top && lam_bottom && g_coef && g_axy && scratch && n_sweeps > 0 && T > 0 &&
                   T <= 65535 && make_gskew(h, w, k),
               "frh_grad_sweeps_adjoint: bad arguments (at most 1024 rows)");
    const size_t state = (size_t)T * 2 * k.plane;
    FRH_GCHECK(n_sweeps == 1 || state_stride >= state, "frh_grad_sweeps_adjoint: state_stride is smaller than one state");
    FRH_GCHECK(lam_top != lam_bottom, "frh_grad_sweeps_adjoint: lam_bottom must not alias lam_top");
    const size_t need = chain_scratch_bytes(n_sweeps, T, k);
    if (scratch_bytes < need)
        return frh_report_error(FRH_EINVAL, "frh_grad_sweeps_adjoint: scratch of %zu bytes needed, %zu given", need,
                                scratch_bytes);
    hipStream_t st = (hipStream_t)stream;
    const int Tpad = (int)chain_tpad(T);
    int *err = (int *)scratch;
    int *flags = (int *)((char *)scratch + 256);
    const int nmax = n_sweeps < kChainMax ? n_sweeps : kChainMax;
    double *lam_mid = (double *)((char *)flags + (size_t)nmax * Tpad * kChainFlagStride * 4);
    const dim3 b(k.hs);
    // launches from the block's last sweeps to its first; a launch's
    // lam_bottom is the next one's lam_top, alternating between the caller's
    // lam_bottom and a scratch plane so that the last launch writes lam_bottom
    const int L = (n_sweeps + nmax - 1) / nmax;
    double *hand = lam_mid + (size_t)(nmax - 1) * state;
    const double *top = lam_top;
    for (int li = 0; li < L; ++li) {
        const int hi = n_sweeps - li * nmax, lo = hi - nmax > 0 ? hi - nmax : 0, n = hi - lo;
        double *bottom = (L - 1 - li) % 2 == 0 ? lam_bottom : hand;
        GChain ch{};
        ch.x0 = lo == 0 ? x0 : xs + (size_t)(lo - 1) * state_stride;
        ch.xs = xs + (size_t)lo * state_stride;
        ch.sstride = state_stride;
        ch.lam_top = top;
        ch.lam_mid = lam_mid;
        ch.lstride = state;
        ch.lam_bottom = bottom;
        ch.flags = flags;
        ch.err = err;
        ch.n = n;
        ch.T = T;
        ch.Tpad = Tpad;
        ch.pub = chain_publish_interval(k.h + k.w - 1);
        const int set = g_coef_set && li == 0 ? 1 : 0;
        if (hipError_t e = launch_clear_flags(flags, n * Tpad, st); e != hipSuccess)
            return grad_fail("chained adjoint flags", e);
        const dim3 g(n * Tpad);
        if (k.hs <= 256)
            hipLaunchKernelGGL((k_gsweep_bwd<4, 4, true>), g, b, 0, st, coef, nullptr, nullptr, nullptr, g_coef,
                               g_axy, nullptr, k, ax, ay, set, ch);
        else if (k.hs <= 512)
            hipLaunchKernelGGL((k_gsweep_bwd<4, 8, true>), g, b, 0, st, coef, nullptr, nullptr, nullptr, g_coef,
                               g_axy, nullptr, k, ax, ay, set, ch);
        else
            hipLaunchKernelGGL((k_gsweep_bwd<2, 16, true>), g, b, 0, st, coef, nullptr, nullptr, nullptr, g_coef,
                               g_axy, nullptr, k, ax, ay, set, ch);
        FRH_GLAUNCH("chained adjoint");
        top = bottom;
    }
    return FRH_OK;
}

int frh_grad_median(const double *x, int N, int h, int w, double *out, const double *g_out, double *g_x,
                    void *stream)
{
    FRH_GCHECK(x && N > 0 && N <= 65535 && h > 0 && h <= 65535 && w > 0, "frh_grad_median: bad arguments");
    hipLaunchKernelGGL(k_gmedian, dim3((w + 63) / 64, h, N), dim3(64), 0, (hipStream_t)stream, x, N, h, w,
                       out, g_out, g_x);
    FRH_GLAUNCH("median");
    return FRH_OK;
}

int frh_grad_warp(const double *img, const double *fill, int fill_per_frame, const double *u,
                  const double *v, int T, int C, int H, int W, double *out, const double *g_out,
                  double *g_img, double *g_u, double *g_v, void *stream)
{
    FRH_GCHECK(img && fill && u && v && T > 0 && T <= 65535 && C > 0 && H > 0 && H <= 65535 && W > 0,
               "frh_grad_warp: bad arguments");
    FRH_GCHECK(!g_out || (g_u && g_v), "frh_grad_warp: g_u and g_v are required with g_out");
    hipLaunchKernelGGL(k_gwarp, dim3((W + 63) / 64, H, T), dim3(64), 0, (hipStream_t)stream, img, fill,
                       fill_per_frame, u, v, T, C, H, W, out, g_out, g_img, g_u, g_v);
    FRH_GLAUNCH("continuous warp");
    return FRH_OK;
}

}  // extern "C"
