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

#include "../../include/flowreg_hip.h"
#include "frh_internal.h"

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

__global__ __launch_bounds__(64) void k_gcoef_fwd(const double *x, const double *u0, const double *J, const double *W, int T,
                            int C, int h, int w, double ax, double ay, const double *a_data, int psi_one,
                            double *coef)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, t = blockIdx.z;
    if (c >= w) return;
    const size_t hw = (size_t)h * w, p = (size_t)r * w + c;
    double Jr[6 * kMaxC], Wr[kMaxC], a[kMaxC], K[6];
    for (int e = 0; e < 6; ++e)
        for (int k = 0; k < C; ++k) Jr[e * C + k] = J[(((size_t)t * 6 + e) * C + k) * hw + p];
    for (int k = 0; k < C; ++k) {
        Wr[k] = W[k * hw + p];
        a[k] = a_data[k];
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
    coef_cell(x[((size_t)t * 2) * hw + p], x[((size_t)t * 2 + 1) * hw + p], ax, ay, a, Jr, Wr, C, psi_one,
              nlr, nud, cu, cv, K);
    for (int e = 0; e < 6; ++e) coef[((size_t)t * 6 + e) * hw + p] = K[e];
}

// Pull-back of the coefficient adjoints gK (T, 6, h, w) through coef_cell:
// gx += dK/dx^T gK (T, 2, h, w), gJ += dK/dJ^T gK, per-cell partial sums
// for ax, ay, a_k in gpar (T, 2 + C, h, w), and (with u0) gu0 += the
// adjoint of the smoothness terms: cu = ax * sx(u) + ay * sy(u) with the
// symmetric neighbour-difference operators sx, sy, so gu0 = ax * sx(gBu) +
// ay * sy(gBu).  Reverse mode by hand:
//   gDu = -omega / Du^2 gK3, gDv = -omega / Dv^2 gK4, gws = gDu + gDv
//   per channel: g_wp = gDu j11 + gDv j22 + gK0 j12 - gK1 j13 - gK2 j23,
//   g_psi = W g_wp; psi = a p, p = (e + 1e-5)^(a - 1), e = max(E, 0):
//   dpsi/da = p (1 + a ln(e + 1e-5)), dpsi/dE = a (a - 1) p / (e + 1e-5)
//   (zero for E < 0); E is quadratic in (x, y) and linear in J.
__global__ __launch_bounds__(64) void k_gcoef_bwd(const double *x, const double *u0, const double *J,
                                                   const double *W, const double *gK, int T, int C, int h,
                                                   int w, double ax, double ay, const double *a_data,
                                                   int psi_one, double *gx, double *gu0, double *gJ,
                                                   double *gpar)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, t = blockIdx.z;
    if (c >= w) return;
    const size_t hw = (size_t)h * w, p = (size_t)r * w + c;
    const double xv = x[((size_t)t * 2) * hw + p], yv = x[((size_t)t * 2 + 1) * hw + p];
    double g[6];
    for (int e = 0; e < 6; ++e) g[e] = gK[((size_t)t * 6 + e) * hw + p];
    // Du, Dv at the refresh (recomputed)
    const double ws = (ax + ax) + (ay + ay);
    double Du = ws, Dv = ws;
    double psi[kMaxC], Jc[6 * kMaxC];
    for (int k = 0; k < C; ++k) {
        for (int e = 0; e < 6; ++e) Jc[e * kMaxC + k] = J[(((size_t)t * 6 + e) * C + k) * hw + p];
        const double j11 = Jc[0 * kMaxC + k], j22 = Jc[1 * kMaxC + k], j33 = Jc[2 * kMaxC + k];
        const double j12 = Jc[3 * kMaxC + k], j13 = Jc[4 * kMaxC + k], j23 = Jc[5 * kMaxC + k];
        psi[k] = psi_one ? 1.0
                         : psi_of(j11 * xv * xv + j22 * yv * yv + 2.0 * (j23 * yv) + 2.0 * (j12 * xv * yv) +
                                      2.0 * (j13 * xv) + j33,
                                  a_data[k]);
        const double wp = W[k * hw + p] * psi[k];
        Du += wp * j11;
        Dv += wp * j22;
    }
    const double gDu = -kOmega / (Du * Du) * g[3], gDv = -kOmega / (Dv * Dv) * g[4];
    const double gws = gDu + gDv;
    double nlr, nud;
    edge_counts(h, w, r, c, nlr, nud);
    double g_ax = 2.0 * gws + nlr * g[5], g_ay = 2.0 * gws + nud * g[5];
    double g_x = 0.0, g_y = 0.0;
    for (int k = 0; k < C; ++k) {
        const double j11 = Jc[0 * kMaxC + k], j22 = Jc[1 * kMaxC + k], j33 = Jc[2 * kMaxC + k];
        const double j12 = Jc[3 * kMaxC + k], j13 = Jc[4 * kMaxC + k], j23 = Jc[5 * kMaxC + k];
        const double Wk = W[k * hw + p], wp = Wk * psi[k];
        double gj[6];
        gj[0] = gDu * wp;
        gj[1] = gDv * wp;
        gj[2] = 0.0;
        gj[3] = g[0] * wp;
        gj[4] = -g[1] * wp;
        gj[5] = -g[2] * wp;
        double g_a = 0.0;
        if (!psi_one) {
            const double g_psi = Wk * (gDu * j11 + gDv * j22 + g[0] * j12 - g[1] * j13 - g[2] * j23);
            const double E = j11 * xv * xv + j22 * yv * yv + 2.0 * (j23 * yv) + 2.0 * (j12 * xv * yv) +
                             2.0 * (j13 * xv) + j33;
            const double a = a_data[k], base = (E < 0 ? 0.0 : E) + 1e-5;
            const double pw = pow(base, a - 1.0);
            g_a = g_psi * pw * (1.0 + a * log(base));
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
        gpar[((size_t)t * (2 + C) + 2 + k) * hw + p] += g_a;
        for (int e = 0; e < 6; ++e) gJ[(((size_t)t * 6 + e) * C + k) * hw + p] += gj[e];
    }
    if (u0) {
        const double *gBu = gK + ((size_t)t * 6 + 1) * hw, *gBv = gK + ((size_t)t * 6 + 2) * hw;
        double sx, sy;
        flow_diffs(u0 + (size_t)t * 2 * hw, h, w, r, c, sx, sy);
        g_ax += g[1] * sx;
        g_ay += g[1] * sy;
        flow_diffs(u0 + ((size_t)t * 2 + 1) * hw, h, w, r, c, sx, sy);
        g_ax += g[2] * sx;
        g_ay += g[2] * sy;
        if (gu0) {
            flow_diffs(gBu, h, w, r, c, sx, sy);
            gu0[((size_t)t * 2) * hw + p] += ax * sx + ay * sy;
            flow_diffs(gBv, h, w, r, c, sx, sy);
            gu0[((size_t)t * 2 + 1) * hw + p] += ax * sx + ay * sy;
        }
    }
    gx[((size_t)t * 2) * hw + p] += g_x;
    gx[((size_t)t * 2 + 1) * hw + p] += g_y;
    gpar[((size_t)t * (2 + C) + 0) * hw + p] += g_ax;
    gpar[((size_t)t * (2 + C) + 1) * hw + p] += g_ay;
}

// --------------------------------------------------------------- sweeps
// One column-major lexicographic SOR sweep (core/level_solver.py:257-364),
// one work-group per frame, lane = row, one diagonal r + c = d per step:
//   nu = Bu + ax (l + r) + ay (d + u) - A12 ov
//   du' = wu nu + (1 - omega + wu b) ou
//   nv = Bv + ax (lv + rv) + ay (dv + uv) - A12 du'
//   dv' = wv nv + (1 - omega + wv b) ov
// (left / up neighbours new, right / down old; neighbours outside the level
// hold the cell's own start value, folded into b).
__global__ void k_gsweep_fwd(const double *coef, const double *xin, double *xout, int h, int w,
                             double ax, double ay)
{
    const int t = blockIdx.x, r = threadIdx.x;
    const size_t hw = (size_t)h * w;
    const double *K = coef + (size_t)t * 6 * hw;
    const double *ui = xin + (size_t)t * 2 * hw, *vi = ui + hw;
    double *uo = xout + (size_t)t * 2 * hw, *vo = uo + hw;
    for (int d = 0; d <= h + w - 2; ++d) {
        const int c = d - r;
        if (r < h && c >= 0 && c < w) {
            const size_t p = (size_t)r * w + c;
            const double A12 = K[p], Bu = K[hw + p], Bv = K[2 * hw + p], wu = K[3 * hw + p],
                         wv = K[4 * hw + p], b = K[5 * hw + p];
            const double lu = c > 0 ? uo[p - 1] : 0.0, uu = r > 0 ? uo[p - w] : 0.0;
            const double ru = c < w - 1 ? ui[p + 1] : 0.0, du = r < h - 1 ? ui[p + w] : 0.0;
            const double lv = c > 0 ? vo[p - 1] : 0.0, uv = r > 0 ? vo[p - w] : 0.0;
            const double rv = c < w - 1 ? vi[p + 1] : 0.0, dv = r < h - 1 ? vi[p + w] : 0.0;
            const double ou = ui[p], ov = vi[p];
            const double nu = Bu + ax * (lu + ru) + ay * (du + uu) - A12 * ov;
            const double un = wu * nu + (1.0 - kOmega + wu * b) * ou;
            const double nv = Bv + ax * (lv + rv) + ay * (dv + uv) - A12 * un;
            uo[p] = un;
            vo[p] = wv * nv + (1.0 - kOmega + wv * b) * ov;
        }
        __syncthreads();
    }
}

// Reverse-mode derivative of k_gsweep_fwd.  lam_out: adjoint of xout;
// writes lam_in (adjoint of xin), adds the coefficient adjoints to gK and
// per-cell partials of ax, ay to gaxy (T, 2, h, w).  Scratch anu, anv, adu,
// adv (T, h, w).  Diagonals run from last to first: a cell's new values
// are consumed only by its right and lower neighbours (diagonal d + 1) and
// by the next sweep (lam_out).
__global__ void k_gsweep_bwd(const double *coef, const double *xin, const double *xout,
                             const double *lam_out, double *lam_in, double *gK, double *gaxy,
                             double *anu, double *anv, double *adu, double *adv, int h, int w,
                             double ax, double ay)
{
    const int t = blockIdx.x, r = threadIdx.x;
    const size_t hw = (size_t)h * w;
    const double *K = coef + (size_t)t * 6 * hw;
    const double *ui = xin + (size_t)t * 2 * hw, *vi = ui + hw;
    const double *uo = xout + (size_t)t * 2 * hw, *vo = uo + hw;
    const double *lu_out = lam_out + (size_t)t * 2 * hw, *lv_out = lu_out + hw;
    double *Anu = anu + (size_t)t * hw, *Anv = anv + (size_t)t * hw;
    double *Adu = adu + (size_t)t * hw, *Adv = adv + (size_t)t * hw;
    double *G = gK + (size_t)t * 6 * hw, *Gxy = gaxy + (size_t)t * 2 * hw;
    for (int d = h + w - 2; d >= 0; --d) {
        const int c = d - r;
        if (r < h && c >= 0 && c < w) {
            const size_t p = (size_t)r * w + c;
            const double A12 = K[p], Bu = K[hw + p], Bv = K[2 * hw + p], wu = K[3 * hw + p],
                         wv = K[4 * hw + p], b = K[5 * hw + p];
            const bool hr = c < w - 1, hd = r < h - 1, hl = c > 0, hu = r > 0;
            const double a_dv = lv_out[p] + (hr ? ax * Anv[p + 1] : 0.0) + (hd ? ay * Anv[p + w] : 0.0);
            const double a_nv = wv * a_dv;
            const double a_du = lu_out[p] + (hr ? ax * Anu[p + 1] : 0.0) + (hd ? ay * Anu[p + w] : 0.0) -
                                A12 * a_nv;
            const double a_nu = wu * a_du;
            Anu[p] = a_nu;
            Anv[p] = a_nv;
            Adu[p] = a_du;
            Adv[p] = a_dv;
            // forward intermediates
            const double lu = hl ? uo[p - 1] : 0.0, uu = hu ? uo[p - w] : 0.0;
            const double ru = hr ? ui[p + 1] : 0.0, du = hd ? ui[p + w] : 0.0;
            const double lv = hl ? vo[p - 1] : 0.0, uv = hu ? vo[p - w] : 0.0;
            const double rv = hr ? vi[p + 1] : 0.0, dv = hd ? vi[p + w] : 0.0;
            const double ou = ui[p], ov = vi[p], un = uo[p];
            const double nu = Bu + ax * (lu + ru) + ay * (du + uu) - A12 * ov;
            const double nv = Bv + ax * (lv + rv) + ay * (dv + uv) - A12 * un;
            G[p] += -ov * a_nu - un * a_nv;                 // A12
            G[hw + p] += a_nu;                             // Bu
            G[2 * hw + p] += a_nv;                         // Bv
            G[3 * hw + p] += (nu + b * ou) * a_du;         // wu
            G[4 * hw + p] += (nv + b * ov) * a_dv;         // wv
            G[5 * hw + p] += wu * ou * a_du + wv * ov * a_dv;  // b
            Gxy[p] += (lu + ru) * a_nu + (lv + rv) * a_nv;
            Gxy[hw + p] += (du + uu) * a_nu + (dv + uv) * a_nv;
        }
        __syncthreads();
    }
    // adjoint of the sweep's input: own old value, and the old values the
    // left and upper neighbours read as their right / lower neighbour
    double *li_u = lam_in + (size_t)t * 2 * hw, *li_v = li_u + hw;
    for (int c = 0; c < w && r < h; ++c) {
        const size_t p = (size_t)r * w + c;
        const double A12 = K[p], wu = K[3 * hw + p], wv = K[4 * hw + p], b = K[5 * hw + p];
        li_u[p] = (1.0 - kOmega + wu * b) * Adu[p] + (c > 0 ? ax * Anu[p - 1] : 0.0) +
                  (r > 0 ? ay * Anu[p - w] : 0.0);
        li_v[p] = (1.0 - kOmega + wv * b) * Adv[p] + (c > 0 ? ax * Anv[p - 1] : 0.0) +
                  (r > 0 ? ay * Anv[p - w] : 0.0) - A12 * Anu[p];
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

__device__ inline double median25_sorted(double *v)
{
    for (int i = 1; i < 25; ++i) {  // insertion sort of 25 values
        double x = v[i];
        int j = i - 1;
        while (j >= 0 && v[j] > x) {
            v[j + 1] = v[j];
            --j;
        }
        v[j + 1] = x;
    }
    return v[12];
}

__global__ __launch_bounds__(64) void k_gmedian(const double *x, int T, int h, int w, double *out, const double *gout,
                          double *gx)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x, r = blockIdx.y, tq = blockIdx.z;
    if (c >= w) return;
    const size_t hw = (size_t)h * w;
    const double *X = x + (size_t)tq * hw;
    double v[25];
    int src[25];
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            const int rr = mirror5(r + i - 2, h), cc = mirror5(c + j - 2, w);
            src[i * 5 + j] = rr * w + cc;
            v[i * 5 + j] = X[rr * w + cc];
        }
    double s[25];
    for (int i = 0; i < 25; ++i) s[i] = v[i];
    const double m = median25_sorted(s);
    if (out) out[(size_t)tq * hw + (size_t)r * w + c] = m;
    if (gx) {
        int k = 0;
        while (k < 24 && v[k] != m) ++k;  // the first window sample equal to the median
        atomicAdd(gx + (size_t)tq * hw + src[k], gout[(size_t)tq * hw + (size_t)r * w + c]);
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

int frh_grad_coef(const double *x, const double *u0, const double *J, const double *W, int T, int C,
                  int h, int w, double ax, double ay, const double *a_data, int psi_one, double *coef,
                  const double *g_coef, double *g_x, double *g_u0, double *g_J, double *g_par,
                  void *stream)
{
    FRH_GCHECK(x && J && W && a_data && T > 0 && C > 0 && C <= kMaxC && h > 0 && w > 0 &&
                   T <= 65535 && h <= 65535, "frh_grad_coef: bad arguments (at most 2 channels)");
    FRH_GCHECK(!g_coef || (g_x && g_J && g_par), "frh_grad_coef: g_x, g_J and g_par are required with g_coef");
    hipStream_t st = (hipStream_t)stream;
    dim3 g((w + 63) / 64, h, T), b(64);
    if (coef) {
        hipLaunchKernelGGL(k_gcoef_fwd, g, b, 0, st, x, u0, J, W, T, C, h, w, ax, ay, a_data, psi_one, coef);
        FRH_GLAUNCH("coefficient forward");
    }
    if (g_coef) {
        hipLaunchKernelGGL(k_gcoef_bwd, g, b, 0, st, x, u0, J, W, g_coef, T, C, h, w, ax, ay, a_data,
                           psi_one, g_x, g_u0, g_J, g_par);
        FRH_GLAUNCH("coefficient backward");
    }
    return FRH_OK;
}

int frh_grad_sweep(const double *coef, const double *x_in, double *x_out, int T, int h, int w,
                   double ax, double ay, void *stream)
{
    FRH_GCHECK(coef && x_in && x_out && T > 0 && T <= 65535 && h > 0 && h <= 1024 && w > 0,
               "frh_grad_sweep: bad arguments (at most 1024 rows)");
    hipLaunchKernelGGL(k_gsweep_fwd, dim3(T), dim3(64 * ((h + 63) / 64)), 0, (hipStream_t)stream, coef,
                       x_in, x_out, h, w, ax, ay);
    FRH_GLAUNCH("sweep forward");
    return FRH_OK;
}

int frh_grad_sweep_adjoint(const double *coef, const double *x_in, const double *x_out,
                           const double *lam_out, double *lam_in, double *g_coef, double *g_axy,
                           double *scratch, int T, int h, int w, double ax, double ay, void *stream)
{
    FRH_GCHECK(coef && x_in && x_out && lam_out && lam_in && g_coef && g_axy && scratch && T > 0 &&
                   T <= 65535 && h > 0 && h <= 1024 && w > 0,
               "frh_grad_sweep_adjoint: bad arguments (at most 1024 rows)");
    const size_t n = (size_t)T * h * w;
    hipLaunchKernelGGL(k_gsweep_bwd, dim3(T), dim3(64 * ((h + 63) / 64)), 0, (hipStream_t)stream, coef,
                       x_in, x_out, lam_out, lam_in, g_coef, g_axy, scratch, scratch + n, scratch + 2 * n,
                       scratch + 3 * n, h, w, ax, ay);
    FRH_GLAUNCH("sweep adjoint");
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
