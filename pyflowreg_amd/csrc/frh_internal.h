// Internal declarations shared by the HIP kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace frh {

// Skewed ("diagonal-major") layout of a level's interior cells used by the
// solver kernels: cell (r, c) of an h x w interior lives at
// (r + c) * hs + r, hs = 64 * ceil(h / 64).  At wavefront step t the SOR
// kernel's lane r touches cell (r, t - r), so every load and store of a step
// is one contiguous run of lanes.  Padding slots (diagonal d, row r with
// r >= h or d - r outside [0, w)) are never read or written.
struct Skew {
    int h, w, hs;
    size_t plane;  // elements per plane: (h + w - 1) * hs
};

inline Skew make_skew(int h, int w)
{
    Skew s;
    s.h = h;
    s.w = w;
    s.hs = 64 * ((h + 63) / 64);
    s.plane = (size_t)(h + w - 1) * (size_t)s.hs;
    return s;
}

__host__ __device__ inline size_t sk(int hs, int r, int c) { return (size_t)(r + c) * hs + r; }

// Stores a formatted message for frh_last_error() and returns code.
int frh_report_error(int code, const char *fmt, ...);

// Data-term exponents of the coefficient formation: a_k, a_k - 1 per
// channel; GNC (core/level_solver.py:372-555): when the stage weight beta is
// set and 0 < a_k < 1, beta < 1, psi_k becomes (1 - beta) + beta * robust.
struct AData {
    double a[8];
    double am1[8];
    int gnc[8];
    double mix, beta;
};

inline AData make_adata(int C, const double *a_data, const double *a_data_minus1, double gnc_beta)
{
    AData ad{};
    for (int k = 0; k < C && k < 8; ++k) {
        ad.a[k] = a_data[k];
        ad.am1[k] = a_data_minus1[k];
        ad.gnc[k] = gnc_beta >= 0 && a_data[k] > 0.0 && a_data[k] < 1.0 && gnc_beta < 1.0;
    }
    ad.mix = 1.0 - gnc_beta;
    ad.beta = gnc_beta;
    return ad;
}

constexpr int kMaxSweepsPerPass = 8;
constexpr int kCoefPlanes = 5;  // Du, Dv, A12, Bu, Bv

// ---- launchers (frh_kernels.hip); all return hipError_t of the launch ----
hipError_t launch_prefilter_v(const float *raw, int T, int H, int W, int C, const double *lo,
                              const double *den, const double *wy, int ry, int sem, double *tmp,
                              hipStream_t st);
hipError_t launch_prefilter_t(const void *raw, int raw_f64, int Tn, int H, int W, int C, const double *lo,
                              const double *den, const double *wt, int rt, double *out, hipStream_t st);
hipError_t launch_prefilter_h(const double *tmp, int T, int H, int W, int C, const double *wx,
                              int rx, int sem, float *out, hipStream_t st);
hipError_t launch_prefilter_h64(const double *tmp, int T, int H, int W, int C, const double *wx,
                                int rx, int sem, double *out, hipStream_t st);
// launch_resize for 4-tap tables (flow upsampling) in one kernel, the same
// results without the intermediate plane
hipError_t launch_resize_fused4(const float *src, int N, int h, int w, float *dst, int oh, int ow,
                                const int32_t *ix, const float *wx, const int32_t *iy, const float *wy,
                                hipStream_t st);
hipError_t launch_resize(const float *src, int N, int h, int w, float *dst, int oh, int ow,
                         const int32_t *ix, const float *wx, const int32_t *iy, const float *wy,
                         int P, float *tmp, hipStream_t st);
// acc64: cv2.remap's 64F path (double sums), else its 32F path (float sums)
hipError_t launch_warp(const float *f2, const float *f1, int f1_per_frame, int T, int C, int H,
                       int W, const float *u, const float *v, size_t uv_frame_stride,
                       double h_col, double h_row, int acc64, float *out, hipStream_t st);
// float64 displacements; images float64 (img_f64, double sums) or float32
hipError_t launch_warp_linear_ex(const void *f2, const void *f1, int img_f64, int acc64, int f1_per_frame,
                                 int T, int C, int H, int W, const double *u, const double *v,
                                 size_t uv_frame_stride, float *out, hipStream_t st);
// np.mean(w, axis=0) of n float32 frames of len values each
hipError_t launch_mean_frames_f32(const float *w, int n, size_t len, float *out, hipStream_t st);
// float32 images holding 16-bit integer samples: float sums, in-bounds
// results rounded half to even and clamped to [lo, hi] (cv2's 16U / 16S
// remap); linear selects INTER_LINEAR, else INTER_CUBIC
hipError_t launch_warp_saturate(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C,
                                int H, int W, const double *u, const double *v, size_t uv_frame_stride,
                                float lo, float hi, float *out, hipStream_t st);
// float32 images holding uint8 samples: cv2's 8U remap (15-bit fixed-point
// weights, integer sums, rounded and saturated to [0, 255])
hipError_t launch_warp_u8(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C, int H,
                          int W, const double *u, const double *v, size_t uv_frame_stride, float *out,
                          hipStream_t st);
hipError_t launch_warp_ex(const void *f2, const void *f1, int img_f64, int acc64, int f1_per_frame, int T,
                          int C, int H, int W, const double *u, const double *v, size_t uv_frame_stride,
                          float *out, hipStream_t st);
hipError_t launch_tensor(int kind, int f2_f32, const float *f1, size_t f1_frame_stride, const float *f2,
                         int T, int C, int h, int w, double hy, double hx, double eps, const float *u,
                         const float *v, size_t uv_frame_stride, double ax, double ay, double *J,
                         double *cuv, size_t frame_stride, const Skew &s, hipStream_t st);
// a_smooth != 1: psi_s over the padded grid at a refresh, nine-coefficient
// assembly, sweeps with per-cell smoothness weights (P <= 2).
hipError_t launch_psi_smooth(const float *u32, const float *v32, size_t fs32, const double *u64,
                             const double *v64, const double *duv, const double *ring, size_t fstride,
                             int T, const Skew &s, int zero, double a_smooth, double gnc_beta, double hx,
                             double hy, double *psi_s, size_t ps_stride, hipStream_t st);
hipError_t launch_assemble_nl(const double *J, const void *Wsk, int w_f32, const float *u32, const float *v32,
                              size_t fs32, const double *u64, const double *v64, const double *psi_s,
                              size_t ps_stride, const double *duv, double *coef, size_t frame_stride, int T,
                              int C, const Skew &s, double ax, double ay, const double *a_data_minus1,
                              const double *a_data, int psi_one, int zero_init, double gnc_beta,
                              hipStream_t st);
hipError_t launch_sweep_nl(int P, const double *coef, double *duv, double *ring, size_t frame_stride, int T,
                           const Skew &s, int zero_init, hipStream_t st);
// sweeps per pass with nine coefficients: 2 up to 512 rows, 1 above (fewer
// registers per lane in the 1024-lane work-groups)
inline int sweep_pmax_nl(int h) { return 64 * ((h + 63) / 64) <= 512 ? 2 : 1; }
hipError_t launch_assemble(const double *J, const double *Wsk, const double *cuv,
                           const double *duv, double *coef, size_t frame_stride, int T, int C,
                           const Skew &s, double ax, double ay, const double *a_data_minus1,
                           const double *a_data, int psi_one, int zero_init, double gnc_beta,
                           hipStream_t st);
// the pyramid solver's assembly: float32 weight planes
hipError_t launch_assemble_w32(const double *J, const float *Wsk, const double *cuv, const double *duv,
                               double *coef, size_t frame_stride, int T, int C, const Skew &s, double ax,
                               double ay, const double *a_data_minus1, const double *a_data, int psi_one,
                               int zero_init, double gnc_beta, hipStream_t st);
hipError_t launch_sweep_chain(int n, int per_stage, const double *coef, double *duv, double *ring,
                              size_t frame_stride, int T, const Skew &s, double ax, double ay, int zero_init,
                              int *flags, int *err, hipStream_t st);
hipError_t launch_sweep_band_chain(int n, int nl, const double *coef, double *duv, double *ring,
                                   size_t frame_stride, int T, const Skew &s, double ax, double ay, int zero_init,
                                   int *flags, int *err, hipStream_t st);
int sweep_chain_slots();
hipError_t launch_sweep(int P, const double *coef, double *duv, double *ring, size_t frame_stride,
                        int T, const Skew &s, double ax, double ay, int zero_init,
                        hipStream_t st);
hipError_t launch_median_add(const double *duv, size_t frame_stride, const Skew &s, int T,
                             int do_median, const float *u, const float *v,
                             size_t uv_frame_stride, float *out_f32, double *out_f64,
                             int out_interleaved, hipStream_t st);
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t st);
// zeroes n progress flags of a chain (one int per kChainFlagStride)
hipError_t launch_clear_flags(int *flags, int n, hipStream_t st);
hipError_t launch_broadcast_f32(const float *src, size_t n, float *dst, size_t dst_stride, int T,
                                hipStream_t st);
hipError_t launch_to_skew_f32f(const float *src, int N, const Skew &s, float *dst, hipStream_t st);
hipError_t launch_to_skew_f32(const float *src, int N, const Skew &s, double *dst,
                              hipStream_t st);
hipError_t launch_prefilter_v64(const double *raw, int T, int H, int W, int C, const double *lo,
                                const double *den, const double *wy, int ry, int sem, double *tmp,
                                hipStream_t st);
hipError_t launch_padded_to_skew_f64(const double *src, int m, int n, int K, const Skew &s,
                                     double *dst, hipStream_t st);
hipError_t launch_skew_to_padded_f64(const double *duv, const double *ring, const Skew &s,
                                     double *du, double *dv, hipStream_t st);
hipError_t launch_median_plane(double *f, int N, int h, int w, double *tmp, hipStream_t st);
hipError_t launch_tensor_rowmajor(int kind, int f1_f32, int f2_f32, const float *f1, const float *f2,
                                  int h, int w, double hy, double hx, double eps, double *J11,
                                  double *J22, double *J33, double *J12, double *J13, double *J23,
                                  hipStream_t st);
hipError_t launch_cuv_padded(const double *u, const double *v, int m, int n, double ax,
                             double ay, const Skew &s, double *cuv, hipStream_t st);

hipError_t launch_pyr_h(const float *src, int N, int H, int W, int Wsum, const int32_t *hoff,
                        const int32_t *hix, const float *hwx, float *tmp, hipStream_t st);
hipError_t launch_pyr_v(const float *tmp, int N, int H, int Wsum, int off, float *dst, int oh,
                        int ow, const int32_t *iy, const float *wy, int P, hipStream_t st);

bool fused_sweep_supported(int C);
hipError_t launch_sweep_fused(int P, const double *J, const double *Wsk, const double *cuv,
                              double *duv, double *ring, size_t frame_stride, int T, int C,
                              const Skew &s, double ax, double ay, const double *a_data_minus1,
                              const double *a_data, int psi_one, int zero_init, double gnc_beta,
                              hipStream_t st);

hipError_t launch_cv_resize(const double *src, int T, int h, int w, int C, double *tmp,
                            double *dst, int oh, int ow, hipStream_t st);

// ------------------------------------------------------------- work lists
// Uneven batches (pairs of different sizes, channel counts and pyramid
// depths) run their level loops in lock-step, aligned at the finest level:
// at level step k every group whose pyramid has reached its level
// li = k - (K - levels) is one entry of a descriptor array, and each stage
// of the level is one launch over all frames of all these groups (work item
// = frame; items of group g are start[g] .. start[g] + T - 1).
constexpr int kWlMaxGroups = 32;

struct WlGroup {
    int T, C, li;        // frames, channels, level index (0 = coarsest)
    int h, w, hs;        // level size and skewed row stride
    size_t plane, fs;    // skewed plane (values), workspace values per frame
    int H, Wsum, xoff;   // full-resolution rows, staged pyramid rows
    int P_img;
    const int32_t *img_iy;
    const float *img_wy;
    int ph, pw;          // previous level (flow upsampling), li > 0
    const int32_t *up_ix, *up_iy;
    const float *up_wx, *up_wy;
    double h_row, h_col, axw, ayw;
    const float *pyr, *pyr1;     // staged rows of the moving / fixed frames
    float *f2l, *f1l, *f2w;      // level images: moving, fixed, warped moving
    float *uvl, *uvn;            // level flow, next level's flow (planes)
    const float *uvtop;          // initial flow at the coarsest level
    const float *Wlev;           // the level's skewed float32 weight planes
    double *work;                // per frame: J (6C), cu/cv, 5 coefficients, du/dv
    int do_median;
    int out_kind;                // 0: uvn planes, 1: float32 / 2: float64 interleaved output, 3: float64 coarse flow
    void *out;                   // output of frame 0 for out_kind 1..3
    AData ad;
};

struct WlStarts {
    int n;                        // groups in the launch
    int start[kWlMaxGroups + 1];  // first item of each group, start[n] = items
};

hipError_t launch_wl_pyr_v(const WlGroup *g, const WlStarts &s, int max_oh, int max_ow, hipStream_t st);
hipError_t launch_wl_flow(const WlGroup *g, const WlStarts &s, int max_h, int max_w, hipStream_t st);
hipError_t launch_wl_warp(const WlGroup *g, const WlStarts &s, int max_h, int max_w, hipStream_t st);
hipError_t launch_wl_tensor(int kind, const WlGroup *g, const WlStarts &s, int max_tiles, double eps,
                            hipStream_t st);
hipError_t launch_wl_assemble(const WlGroup *g, const WlStarts &s, int max_diag, int max_hs, int psi_one,
                              int zero_init, hipStream_t st);
// form1024: 1024-lane work-groups (levels of 513 .. 1024 rows, P <= 2)
hipError_t launch_wl_sweep(int P, int form1024, const WlGroup *g, const WlStarts &s, int zero_init,
                           hipStream_t st);
hipError_t launch_wl_median(const WlGroup *g, const WlStarts &s, int max_blocks, hipStream_t st);

}  // namespace frh
