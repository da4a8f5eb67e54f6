/*
 * flowreg_hip.h -- C ABI of libflowreg_hip.so, the MI355X (gfx950) backend
 * for pyflowreg's variational optical-flow path.
 *
 * All array arguments are raw pointers.  Unless stated otherwise they point
 * to device memory (HBM) and are C-contiguous; `stream` is a hipStream_t
 * (NULL = default stream).  Every entry point returns FRH_OK (0) or a
 * negative error code; frh_last_error() then returns a message that names
 * the failing check or kernel.  No entry point falls back to the CPU.
 *
 * Layout conventions
 *   image planes : (N, H, W) float32, one plane per (frame, channel)
 *   flows        : (H, W) float32 planes, u = x (columns), v = y (rows)
 *   level solver : (m, n, K) float64 interleaved, like the reference's
 *                  numpy arrays (m = rows + 2, n = cols + 2, K channels)
 *
 * Reference citations are pyflowreg source paths (file:line).
 */
#ifndef FLOWREG_HIP_H
#define FLOWREG_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FRH_OK 0
#define FRH_EINVAL (-1)   /* bad argument (shape, size, pointer) */
#define FRH_EHIP (-2)     /* a HIP runtime call or kernel launch failed */
#define FRH_ENOMEM (-3)   /* device allocation failed */
#define FRH_EUNSUP (-4)   /* valid request this build does not implement */

const char *frh_last_error(void);
/* Incremented whenever entry points are added or change; 3 added
 * frh_warp_int16 and frh_mean_frames_f32, 4 frh_warp_u8; 5 moved every
 * stage's scratch to the caller (scratch arguments and *_scratch_bytes
 * queries; frh_prefilter / frh_prefilter_f64 folded into frh_prefilter_ex);
 * 6 adds frh_prefilter_noalloc. */
int frh_abi_version(void);

/* ------------------------------------------------------------ host only */

/* util/resize_util.py:194-213 (_fill_tables_fused_gauss_cubic_reflect).
 * Host memory.  g: the 2R+1 float32 Gaussian taps of
 * util/resize_util.py:216-225 (R = 0: g = {1}).  idx, wt: (out_len, 2R+4). */
int frh_fill_resize_tables(int in_len, int out_len, int R, const float *g,
                           int32_t *idx, float *wt);

/* core/warping.py:237-277 (warpingDepth). */
int frh_warping_depth(double eta, int levels, int m, int n);

/* ------------------------------------------------------------ stages */

/* Scratch.  The stage entry points below that need device memory beyond
 * their inputs and outputs take it from the caller (scratch, at least
 * scratch_bytes as returned by the matching *_scratch_bytes query; a
 * smaller or missing buffer is rejected with FRH_EINVAL): no stage entry
 * point allocates.  The scratch is used on `stream` only, so a buffer may be
 * reused by the next call on the same stream. */

/* Bytes of scratch frh_prefilter_ex (rt = 0) and frh_prefilter_3d need for
 * T frames of H x W x C: the float64 planes between the two Gaussian
 * passes, and with a temporal kernel (rt > 0) the frames after it. */
size_t frh_prefilter_scratch_bytes(int T, int H, int W, int C, int rt);

/* util/image_processing.py:12-98 (normalize) followed by :101-165
 * (apply_gaussian_filter, mode "reflect"), then the executors' float32 cast
 * (motion_correction/parallelization/base.py:63-73).
 * raw: (T, H, W, C), float64 when raw_f64 != 0 (the reference frame is the
 * float64 mean of raw frames, motion_correction/compensate_recording.py:
 * 491-519), else float32.  lo[c], den[c]: normalisation offset and
 * denominator (max - min + eps) per channel, float64, host memory.
 * wy (2*ry+1), wx (2*rx+1): Gaussian weights, float64, host memory, the
 * values scipy.ndimage uses.  out: (T, C, H, W) planes, float32, or with
 * out_f64 = 1 float64: the preprocessed frames before the executors'
 * float32 cast, as update_reference (compensate_recording.py:680-713) and
 * FlowRegLive (flow_reg_live.py:216-323) use them.  Arithmetic is float64
 * in scipy's summation order (vertical pass, then horizontal).
 * semantics: 0 for that float64 arithmetic, or flags for a float32 frame
 * normalised by its own float32 range (util/image_processing.py:12-98 keeps
 * numpy's float32 there, FlowRegLive's reference): FRH_PRE_NORM32,
 * (x - lo) / den in float32; FRH_PRE_FILT32, each Gaussian pass stored as
 * float32, as scipy's gaussian_filter does for a float32 array. */
#define FRH_PRE_NORM32 1
#define FRH_PRE_FILT32 2
int frh_prefilter_ex(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                     const double *den, const double *wy, int ry, const double *wx, int rx,
                     void *out, int out_f64, int semantics, void *scratch, size_t scratch_bytes,
                     void *stream);

/* frh_prefilter_ex without a scratch argument and without any device
 * allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, tight
 * per-frame serving loops such as a live registration, harnesses that
 * forbid hipMalloc).  Its scratch is a static device array of
 * frh_prefilter_noalloc_max_bytes() bytes (32 MiB) that each GPU holds until
 * the process ends; calls needing more (T*C*H*W float64 values) are rejected
 * with FRH_EINVAL.  Every call on one GPU shares that GPU's array: calls that
 * overlap in time on one GPU (different streams, or threads) race.  The
 * stream must belong to the current device.  Results equal frh_prefilter_ex's
 * bit for bit. */
int frh_prefilter_noalloc(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                          const double *den, const double *wy, int ry, const double *wx, int rx,
                          void *out, int out_f64, int semantics, void *stream);
size_t frh_prefilter_noalloc_max_bytes(void);

/* apply_gaussian_filter's 4-D branch with a temporal kernel
 * (util/image_processing.py:147-161; scipy's gaussian_filter filters the
 * frame axis first, then rows, then columns): raw (T, H, W, C) normalised
 * with (lo, den) per channel, correlated along T with wt (2 rt + 1 taps,
 * mode "reflect" over the T frames of the call), then the row and column
 * passes of frh_prefilter_ex (float64 semantics).  rt = 0 is frh_prefilter_ex
 * with semantics 0. */
int frh_prefilter_3d(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                     const double *den, const double *wt, int rt, const double *wy, int ry,
                     const double *wx, int rx, void *out, int out_f64, void *scratch,
                     size_t scratch_bytes, void *stream);

/* util/resize_util.py:116-141,232-254 (imresize_fused_gauss_cubic).
 * src: (N, h, w) planes; dst: (N, oh, ow); tmp: (N, h, ow) scratch.
 * ix, wx: (ow, P) and iy, wy: (oh, P) device tables from
 * frh_fill_resize_tables.  float32 products, float64 sums. */
int frh_resize(const float *src, int N, int h, int w, float *dst, int oh, int ow,
               const int32_t *ix, const float *wx, const int32_t *iy, const float *wy, int P,
               float *tmp, void *stream);

/* core/warping.py:160-234 (imregister_wrapper, cubic) of float32 images
 * (cv2.remap's 32F path: float sums).  f2: (T, C, H, W) planes to warp; f1:
 * (C, H, W) fill for out-of-bounds pixels, or (T, C, H, W) when
 * f1_per_frame != 0.  u, v: (T, H, W) float32 displacements, divided by
 * h_col / h_row (float64) before use.  out: (T, C, H, W).  OpenCV
 * INTER_CUBIC + BORDER_REPLICATE semantics with 1/32-px map quantization. */
int frh_warp_cubic(const float *f2, const float *f1, int f1_per_frame, int T, int C, int H,
                   int W, const float *u, const float *v, double h_col, double h_row,
                   float *out, void *stream);

/* The same warp for float64 displacements already divided by the grid
 * spacing (the reference's imregister_wrapper(f2, u / h_col, v / h_row, f1)
 * with u of any dtype: the map is float32(x + u) formed in float64), with the
 * image type the reference passes: float64 images take cv2.remap's 64F path
 * (float weights, double sums; acc_f64 = 1) -- the level images of
 * get_displacement are float64 -- and float32 images its 32F path
 * (acc_f64 = 0).  f2, f1 are float64 when img_f64 != 0 (then acc_f64 must be
 * 1), else float32 (float64 values exactly representable in float32 may be
 * passed as float32 with acc_f64 = 1).  The result and the out-of-bounds fill
 * are rounded to float32, as the reference's float32 output array does. */
int frh_warp_cubic_ex(const void *f2, const void *f1, int img_f64, int acc_f64, int f1_per_frame,
                      int T, int C, int H, int W, const double *u, const double *v, float *out,
                      void *stream);

/* imregister_wrapper(f2, u, v, f1, interpolation_method="linear")
 * (core/warping.py:160-234 with cv2.INTER_LINEAR): the same arguments and
 * image-type rules as frh_warp_cubic_ex, bilinear taps on the 1/32-px grid
 * with BORDER_REPLICATE.  Used by the cross-correlation prealignment
 * (motion_correction/parallelization/sequential.py:97-130). */
int frh_warp_linear_ex(const void *f2, const void *f1, int img_f64, int acc_f64, int f1_per_frame,
                       int T, int C, int H, int W, const double *u, const double *v, float *out,
                       void *stream);

/* imregister_wrapper(f2, u, v, f1, ...) with uint16 (is_signed = 0) or int16
 * (is_signed = 1) raw frames f2 (core/warping.py:219-227: cv2.remap keeps
 * the integer depth).  OpenCV's 16U / 16S remap sums the integer samples with
 * its float weights, like its 32F path, and stores saturate_cast<T>(sum):
 * rounded half to even and clamped to the type's range.  f2 holds the samples
 * converted to float32 (exact); f1, the out-of-bounds fill, is float32 and
 * stored unrounded (the reference writes it into its float32 output array).
 * linear = 1: INTER_LINEAR, else INTER_CUBIC; other arguments as
 * frh_warp_cubic_ex. */
int frh_warp_int16(int linear, int is_signed, const float *f2, const float *f1, int f1_per_frame, int T,
                   int C, int H, int W, const double *u, const double *v, float *out, void *stream);

/* imregister_wrapper(f2, u, v, f1, ...) with uint8 raw frames f2
 * (core/warping.py:219-227: cv2.remap keeps the 8-bit depth).  OpenCV's 8U
 * remap weighs the taps with its 15-bit fixed-point table (the float weight
 * products times 2^15, rounded, their sum fixed up to 2^15 on one central
 * cell), sums the integer samples exactly and stores
 * saturate_cast<uchar>((sum + 2^14) >> 15).  f2 holds the samples converted
 * to float32 (exact); f1, the out-of-bounds fill, is float32 and stored as
 * it is.  linear = 1: INTER_LINEAR, else INTER_CUBIC; other arguments as
 * frh_warp_cubic_ex. */
int frh_warp_u8(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C, int H, int W,
                const double *u, const double *v, float *out, void *stream);

/* Data terms (core/optical_flow.py:446-473, _resolve_motion_tensor_func). */
#define FRH_TENSOR_GC 0   /* "gc" / "gradient": get_motion_tensor_gc, :73-160 */
#define FRH_TENSOR_GRAY 1 /* "gray" / "brightness": get_motion_tensor_gray, :163-228 */
#define FRH_TENSOR_CS 2   /* "cs" / "census": get_motion_tensor_cs, :231-345 */

/* get_motion_tensor_{gc,gray,cs}(f1, f2, hy, hx[, eps]) for one channel.
 * f1, f2: (h, w) float32-valued images.  f1_f32 / f2_f32: the caller's
 * array is float32 (numpy then evaluates that image's derivatives in
 * float32, as for the warped image below the top level); 0 = float64.
 * eps: census smoothing width (ignored by gc and gray; the reference's
 * default is 0.1 / 255).  J11..J23: (h+2, w+2) float64, zero border. */
int frh_motion_tensor(int kind, const float *f1, const float *f2, int h, int w, double hy,
                      double hx, double eps, int f1_f32, int f2_f32, double *J11, double *J22,
                      double *J33, double *J12, double *J13, double *J23, void *stream);

/* core/optical_flow.py:348-411 (level_solver) -> core/level_solver.py:158-369
 * (compute_flow), or :372-555 (compute_flow_gnc) when gnc_beta >= 0 (the
 * reference's gnc_beta=None is any negative value).  Same positional
 * meaning as the reference: J*, weight (m, n, K) float64; u, v (m, n)
 * float64; a_data (K) float64 host memory; hx scales column differences,
 * hy rows.  du, dv: (m, n) float64 outputs (interior updated, ring = the
 * last boundary copy).  Sweeps run in the reference's column-major
 * lexicographic order (a diagonal wavefront on the GPU).  a_smooth != 1
 * refreshes the smoothness weights psi_s (nonlinearity_smoothness_2d,
 * :69-155) with the data term's psi and weights each edge by
 * 0.5 (psi_s[c] + psi_s[nb]) alpha / h^2 (:277-316). */
int frh_level_solver(const double *J11, const double *J22, const double *J33,
                     const double *J12, const double *J13, const double *J23,
                     const double *weight, const double *u, const double *v, int m, int n,
                     int K, double alpha_x, double alpha_y, int iterations, int update_lag,
                     const double *a_data, double a_smooth, double hx, double hy,
                     double gnc_beta, double *du, double *dv, void *scratch, size_t scratch_bytes,
                     void *stream);
/* Bytes of scratch frh_level_solver needs for (m, n, K) arrays: the skewed
 * copies of J, the weights and the initial flow, the coefficients, du / dv
 * and the boundary ring (and psi_s when a_smooth != 1). */
size_t frh_level_solver_scratch_bytes(int m, int n, int K, double a_smooth);

/* cv2.resize(w, (ow, oh), interpolation=cv2.INTER_CUBIC) of float64 flows
 * (core/optical_flow.py:615-616, only when min_level > 0), restated from
 * OpenCV's public behaviour.  src: (T, h, w, C), dst: (T, oh, ow, C). */
int frh_resize_cubic_cv(const double *src, int T, int h, int w, int C, double *dst, int oh,
                        int ow, void *scratch, size_t scratch_bytes, void *stream);
/* Bytes of scratch of frh_resize_cubic_cv: the horizontal pass's output. */
size_t frh_resize_cubic_cv_scratch_bytes(int T, int h, int ow, int C);

/* scipy.ndimage.median_filter(size=(5, 5), mode="mirror") as called at
 * core/optical_flow.py:607-609.  f: (N, h, w) float64 planes, in place. */
int frh_median5_mirror(double *f, int N, int h, int w, void *scratch, size_t scratch_bytes,
                       void *stream);
/* Bytes of scratch of frh_median5_mirror: a copy of the planes. */
size_t frh_median5_mirror_scratch_bytes(int N, int h, int w);

/* np.mean(w, axis=0) of n float32 frames of len values each, w C-contiguous
 * (n, len): the rows summed in order in float32, then divided by n
 * (motion_correction/compensate_recording.py:674,801-806, the w_init
 * means of the batch loop).  out holds len float32 values. */
int frh_mean_frames_f32(const float *w, int n, size_t len, float *out, void *stream);

/* ----------------------------------------------------- whole pyramid */

/* One pyramid level as planned by the host (core/optical_flow.py:514-577).
 * Tables are host pointers; they are copied to the device at create. */
typedef struct {
    int h, w;                 /* level size (rows, cols) */
    double h_row, h_col;      /* m / h, n / w */
    double alpha_x, alpha_y;  /* alpha scaled by eta^(-i/2) (1 at min_level) */
    int P_img;                /* taps of the full-res -> level resize */
    const int32_t *img_ix;    /* (w, P_img) */
    const float *img_wx;
    const int32_t *img_iy;    /* (h, P_img) */
    const float *img_wy;
    int P_up;                 /* taps of the coarser -> this level resize (0 at top) */
    const int32_t *up_ix;     /* (w, P_up) */
    const float *up_wx;
    const int32_t *up_iy;     /* (h, P_up) */
    const float *up_wy;
} frh_level_desc;

typedef struct {
    int H, W, C;              /* full-resolution frame geometry */
    int n_levels;             /* levels[0] = coarsest (top) ... last = finest */
    const frh_level_desc *levels;
    int iterations, update_lag;
    double a_smooth;
    const double *a_data;     /* (C) host */
    int max_batch;            /* largest T passed to frh_solver_run */
    int const_assumption;     /* FRH_TENSOR_GC, _GRAY or _CS */
} frh_solver_desc;

typedef struct frh_solver frh_solver;

/* Allocates every device workspace for up to max_batch frames. */
int frh_solver_create(frh_solver **out, const frh_solver_desc *desc);
void frh_solver_destroy(frh_solver *s);

/* Per-recording inputs shared by every frame of a batch
 * (core/optical_flow.py:519,531,565): fixed (C, H, W) float32 planes (or
 * NULL when only frh_solver_run_pairs follows), weight (C, H, W) float32
 * planes (float32-cast weight map), uv (2, H, W) float32 initial flow or
 * NULL (zero). */
int frh_solver_set_reference(frh_solver *s, const float *fixed, const float *weight,
                             const float *uv, void *stream);

/* core/optical_flow.py:773-906 (get_displacement) for T moving frames
 * (T, C, H, W) float32 against the reference: w_out (T, H, W, 2) float64
 * or float32 (out_f64 = 0), [..., 0] = u, [..., 1] = v. */
int frh_solver_run(frh_solver *s, const float *moving, int T, void *w_out, int out_f64,
                   void *stream);

/* get_displacement for T independent pairs of one size: fixed[t] against
 * moving[t], both (T, C, H, W) float32 planes, sharing the weight and the
 * initial flow of frh_solver_set_reference.  Each pair's result equals a
 * separate get_displacement call (uneven batches group pairs by size). */
int frh_solver_run_pairs(frh_solver *s, const float *fixed, const float *moving, int T,
                         void *w_out, int out_f64, void *stream);

/* Work list of uneven pairs (one launch per stage and level step for all
 * frames of all solvers): solvers[0..n) (n <= 32, one per frame size and
 * channel count; equal iterations, update_lag and data term, a_smooth = 1,
 * at most 4 channels, levels of at most 1024 rows, every solver's weight and
 * initial flow set by frh_solver_set_reference) run the level loop of
 * frh_solver_run_pairs in lock-step, aligned at the finest level.
 * frh_worklist_run takes, per solver k, T[k] pairs fixed[k], moving[k]
 * ((T[k], C, H, W) float32 planes) and writes w_out[k] ((T[k], H, W, 2),
 * float64 or float32).  Each pair's result equals its own get_displacement
 * call (core/optical_flow.py:773-929) bit for bit; the batching replaces
 * the reference's per-pair calls of its executors
 * (motion_correction/parallelization/sequential.py:141-145).  The solvers'
 * workspaces are used, so a solver runs in one work list or alone at a
 * time.  The passes follow the default plan of at most 5 sweeps per pass
 * (levels of up to 512 rows) or 2 (513 .. 1024 rows); limits set with
 * frh_solver_set_pass_limit do not apply.  The work list keeps the solver
 * handles; destroy it before them. */
typedef struct frh_worklist frh_worklist;
int frh_worklist_create(frh_worklist **out, frh_solver *const *solvers, int n);
void frh_worklist_destroy(frh_worklist *wl);
int frh_worklist_run(frh_worklist *wl, const float *const *fixed, const float *const *moving, const int *T,
                     void *const *w_out, int out_f64, void *stream);

/* Kernel-time accounting of the last frh_solver_run, measured with HIP
 * events on the run's stream: ms[k] is the summed duration and counts[k]
 * the number of launch groups of category k: 0 = whole run, 1 = SOR
 * sweeps, 2 = coefficient assembly, 3 = motion tensor, 4 = resize (pyramid
 * + flow), 5 = warp, 6 = median/add.  Requires frh_solver_profile(s, 1)
 * before the run (it synchronizes the stream); returns the count written. */
int frh_solver_profile(frh_solver *s, int enable);

/* enable != 0: frh_solver_run / frh_solver_run_pairs replay the level loop
 * as a hipGraph captured on the first call with a given (T, out_f64,
 * per-frame references), with the frames and flows copied through
 * solver-owned staging buffers (device-to-device, on the caller's stream).
 * Results equal the eager launches bit for bit; profiling runs eagerly.
 * No counterpart in the reference (its solvers are CPU or per-call GPU
 * code); it replaces per-kernel launch overhead for small batches. */
int frh_solver_set_graph(frh_solver *s, int enable);
int frh_solver_timings(frh_solver *s, double *ms, int *counts, int n);

/* enable != 0 (the default): when a pass's sweeps x frames (x 512-row
 * bands on levels of more than 512 rows) fit one work-group per compute
 * unit, the pass runs as a chain of work-groups, one per sweep (and band)
 * and frame, each consuming its predecessor's results as they are
 * published; bit-identical to the one-work-group-per-frame passes, and
 * faster when the batch leaves compute units idle.  Disable it when other
 * work shares the device (several solvers on concurrent streams).  No
 * counterpart in the reference. */
int frh_solver_set_chains(frh_solver *s, int enable);

/* Tuning knobs of the SOR wavefront (no counterpart in the reference, whose
 * solvers have no launch geometry): each level's sweeps run as wavefront
 * passes of at most pass_limit sweeps (between coefficient refreshes); more
 * sweeps per pass read and write the state fewer times but hold more
 * registers.  The result does not depend on the limit.  level_info returns
 * a level's size, current and largest allowed limit, and its sweep +
 * assembly time (ms) in the last profiled run. */
int frh_solver_level_info(frh_solver *s, int level, int *h, int *w, int *pass_limit, int *pass_limit_max,
                          double *last_ms);
int frh_solver_set_pass_limit(frh_solver *s, int level, int pass_limit);

/* ------------------------------------------- differentiable registration */
/* Forward and backward (reverse-mode) kernels of one pass of
 * get_displacement (core/optical_flow.py:476-617: gradient-constancy data
 * term, core/level_solver.py:158-369, a_smooth = 1) and of a continuous
 * bicubic warp (core/warping.py:160-234 without the 1/32-px map rounding),
 * all float64, for gradients of a registration loss w.r.t. alpha, a_data
 * and the moving frames.  The reference has no gradients, so these entry
 * points have no counterparts there; their forward operations are the
 * reference's.
 *   images (T, C, h, w) and J (T, 6, C, h, w) row-major;
 *   solver states (T, plane, 2) = (du, dv), coefficients (T, plane, 6) =
 *   (A12, Bu, Bv, omega/Du, omega/Dv, b) and their adjoints in the skewed
 *   layout with each cell's values interleaved: cell (r, c) at slot
 *   (r + c) * hs + r, hs = 64 * ceil(h / 64), plane = (h + w - 1) * hs
 *   slots (frh_grad_skew_plane); padding slots are never read.  At most 2
 *   channels and 1024 rows per level. */

/* Slots per skewed plane of an h x w level. */
int frh_grad_skew_plane(int h, int w, size_t *plane);

/* N groups of K planes (N, K, h, w) row-major -> (N, plane, K) skewed with
 * the K values of a cell interleaved (to_skew = 1, padding slots set to 0),
 * or back (to_skew = 0). */
int frh_grad_skew(const double *src, int N, int K, int h, int w, double *dst, int to_skew, void *stream);

/* get_motion_tensor_gc of fixed (C, h, w) and moving (T, C, h, w): J = f
 * (when J != NULL); g_moving += dJ/dmoving^T gJ (when gJ != NULL). */
int frh_grad_tensor(const double *fixed, const double *moving, int T, int C, int h, int w,
                    double hy, double hx, double *J, const double *gJ, double *g_moving,
                    void *stream);

/* Lag-block coefficients from the state x (skewed) at a psi refresh (psi =
 * 1 when psi_one), the level's initial flow u0 (T, 2, h, w; NULL = zero), J,
 * weight W (C, h, w), ax = alpha_x / hx^2, ay = alpha_y / hy^2 and a_data
 * (C, device): coef (skewed, when != NULL).  With g_coef (skewed), adds the
 * pull-backs g_x (skewed), g_u0 (row-major, when u0 and g_u0 are not NULL),
 * g_J and partial sums g_par (T, 2 + C, plane) of ax, ay, a_k over tiles of
 * 16 x 16 cells (tile (i, j) -- rows 16 i.., columns 16 j.. -- at element
 * i ceil(w / 16) + j of its plane; the rest untouched: sum a plane for the
 * frame's total). */
int frh_grad_coef(const double *x, const double *u0, const double *J, const double *W, int T, int C,
                  int h, int w, double ax, double ay, const double *a_data, int psi_one, double *coef,
                  const double *g_coef, double *g_x, double *g_u0, double *g_J, double *g_par,
                  void *stream);

/* One column-major lexicographic SOR sweep x_in -> x_out (skewed). */
int frh_grad_sweep(const double *coef, const double *x_in, double *x_out, int T, int h, int w,
                   double ax, double ay, void *stream);

/* n_sweeps consecutive frh_grad_sweep calls with one coefficient set (the
 * sweeps of a lag block); the results equal the single sweeps'.
 * write_all != 0: the state after sweep j goes to x_out + j out_stride
 * (values; x_in must not overlap them); else only the last state, to x_out
 * (which may equal x_in).  Without scratch (NULL): pipelined in one
 * wavefront pass per at most 5 (levels of up to 256 rows) or 4 (512) sweeps
 * (single sweeps above 512 rows), one work-group per frame.  With scratch
 * (frh_grad_sweeps_scratch_bytes bytes): one launch per at most 32 sweeps
 * as a chain of work-groups per frame, one per sweep, each reading its
 * predecessor's output as it is published (as frh_grad_sweeps_adjoint; the
 * first int of scratch is the failure word) -- for batches that leave most
 * compute units idle. */
size_t frh_grad_sweeps_scratch_bytes(int n_sweeps, int T);
int frh_grad_sweeps(const double *coef, const double *x_in, double *x_out, size_t out_stride, int n_sweeps,
                    int write_all, int T, int h, int w, double ax, double ay, void *scratch, size_t scratch_bytes,
                    void *stream);

/* Exact adjoint of frh_grad_sweep (all skewed): lam_in = (dx_out/dx_in)^T
 * lam_out; g_coef += (dx_out/dcoef)^T lam_out (g_coef = ... when
 * g_coef_set != 0: the first adjoint sweep of a lag block needs no cleared
 * buffer); g_axy (T, 2, plane) += partials of ax, ay (one per row r, at
 * element r of each plane). */
int frh_grad_sweep_adjoint(const double *coef, const double *x_in, const double *x_out,
                           const double *lam_out, double *lam_in, double *g_coef, int g_coef_set, double *g_axy,
                           int T, int h, int w, double ax, double ay, void *stream);

/* The adjoints of a lag block's n_sweeps sweeps with one coefficient set
 * (frh_grad_sweeps) in one launch per at most 32 sweeps: a chain of
 * work-groups per frame, one per sweep, each passing its lam_in to the next
 * through global memory as it goes (agent-scope release / acquire of
 * per-stage progress flags); bit-identical to frh_grad_sweep_adjoint
 * called from the block's last sweep to its first.  x0: the block's start
 * state; the state after sweep j at xs + j state_stride (values); lam_top:
 * adjoint of the last state; lam_bottom (output, not lam_top): adjoint of
 * x0; g_coef and g_axy as in frh_grad_sweep_adjoint (g_coef_set: the
 * block's last sweep stores).  scratch: frh_grad_sweeps_adjoint_scratch_bytes
 * bytes; its first int is set nonzero when a stage gave up waiting for its
 * predecessor (bounded waits: the launch ends, its results are invalid) --
 * the caller zeroes it once and checks it after the launches.  Relies on
 * work-groups being dispatched in index order (a stage's predecessor has the
 * lower index). */
size_t frh_grad_sweeps_adjoint_scratch_bytes(int n_sweeps, int T, int h, int w);
int frh_grad_sweeps_adjoint(const double *coef, const double *x0, const double *xs, size_t state_stride,
                            int n_sweeps, const double *lam_top, double *lam_bottom, double *g_coef, int g_coef_set,
                            double *g_axy, int T, int h, int w, double ax, double ay, void *scratch,
                            size_t scratch_bytes, void *stream);

/* 5x5 mirror median of N (h, w) planes: out (when != NULL); g_x += the
 * output gradient g_out routed to the selected samples (when g_x != NULL). */
int frh_grad_median(const double *x, int N, int h, int w, double *out, const double *g_out,
                    double *g_x, void *stream);

/* Continuous warp of img (T, C, H, W) by (u, v) (T, H, W) with fill
 * (C, H, W) or (T, C, H, W) outside: out (when != NULL); with g_out:
 * g_img += (dout/dimg)^T g_out (when g_img != NULL) and g_u, g_v = the
 * flow gradients. */
int frh_grad_warp(const double *img, const double *fill, int fill_per_frame, const double *u,
                  const double *v, int T, int C, int H, int W, double *out, const double *g_out,
                  double *g_img, double *g_u, double *g_v, void *stream);

#ifdef __cplusplus
}
#endif

#endif
