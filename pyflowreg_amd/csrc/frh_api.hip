// C ABI of libflowreg_hip.so: argument checking, host-side planning, device
// workspaces and the pyramid level loop that drives the kernels.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flowreg_hip.h"
#include "frh_internal.h"
#include "frh_device.h"

using namespace frh;

// the census smoothing width the reference's get_motion_tensor_cs uses when
// called from the level loop (core/optical_flow.py:287-288)
static constexpr double kCensusEps = 0.1 / 255.0;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

#define FRH_HIP(call, what)                                                                   \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess) return fail(FRH_EHIP, "%s: %s", what, hipGetErrorString(e_)); \
    } while (0)

#define FRH_REQ(cond, ...)                                   \
    do {                                                     \
        if (!(cond)) return fail(FRH_EINVAL, __VA_ARGS__);   \
    } while (0)

hipStream_t as_stream(void *s) { return (hipStream_t)s; }

double py_round(double x) { return std::nearbyint(x); }

// util/resize_util.py:68-76, A = -0.75
double keys_cubic(double x)
{
    const double A = -0.75;
    double ax = std::fabs(x);
    if (ax < 1.0) return (A + 2.0) * ax * ax * ax - (A + 3.0) * ax * ax + 1.0;
    if (ax < 2.0) return A * ax * ax * ax - 5.0 * A * ax * ax + 8.0 * A * ax - 4.0 * A;
    return 0.0;
}

int reflect_index(int j, int n)
{
    if (n <= 1) return 0;
    while (j < 0 || j >= n) j = j < 0 ? -j - 1 : 2 * n - 1 - j;
    return j;
}

// Pass schedule of compute_flow (core/level_solver.py:229-258): psi is
// refreshed before sweep `it` when (it + 1) % update_lag == 0; between
// refreshes the coefficients are constant, so a segment of sweeps runs as one
// or more wavefront passes of at most pmax sweeps.
struct Pass {
    int nsweeps;
    bool assemble;  // form coefficients before this pass
    bool psi_one;   // psi = 1 (no refresh has happened yet)
};

// Passes can form their coefficients on entry (fused) when every lag block
// fits in one pass, i.e. no pass continues a block started by another.
bool passes_fusable(const std::vector<Pass> &ps)
{
    for (const Pass &p : ps)
        if (!p.assemble) return false;
    return true;
}

std::vector<Pass> plan_passes(int iterations, int lag, int pmax)
{
    std::vector<int> starts;
    for (int it = 0; it < iterations; ++it)
        if (it == 0 || (it + 1) % lag == 0) starts.push_back(it);
    std::vector<Pass> out;
    for (size_t k = 0; k < starts.size(); ++k) {
        int b = starts[k], e = k + 1 < starts.size() ? starts[k + 1] : iterations;
        bool refresh = ((b + 1) % lag == 0);
        bool first = true;
        for (int s = b; s < e; s += pmax) {
            int n = e - s < pmax ? e - s : pmax;
            out.push_back({n, first, !refresh});
            first = false;
        }
    }
    return out;
}

// Sweeps per wavefront pass: 5 for 512-lane work-groups, 2 for 1024-lane
// ones, 1 for levels over 1024 rows (band mode, one launch per 1024 rows).
int sweep_pmax(int h)
{
    const int hs = 64 * ((h + 63) / 64);
    return hs <= 512 ? 5 : (hs <= 1024 ? 2 : 1);
}

template <typename T>
int dev_alloc(T **p, size_t n, const char *what)
{
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) return fail(FRH_ENOMEM, "hipMalloc(%s, %zu bytes): %s", what, sizeof(T) * n, hipGetErrorString(e));
    return FRH_OK;
}

template <typename T>
int upload(T **dev, const T *host, size_t n, const char *what)
{
    int rc = dev_alloc(dev, n, what);
    if (rc) return rc;
    FRH_HIP(hipMemcpy(*dev, host, sizeof(T) * n, hipMemcpyHostToDevice), what);
    return FRH_OK;
}

}  // namespace

// error reporting shared with the other translation units of the library
int frh::frh_report_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

// ------------------------------------------------ no-allocation prefilter
// frh_prefilter_noalloc takes its scratch from this file-scope device array
// instead of a caller buffer: every GPU the library runs on holds its own
// copy (32 MiB) until the process ends.  Host code never passes the array's
// name to a kernel (on the host it is only a stand-in); it uses the device
// address hipGetSymbolAddress returns, looked up once per device with that
// device current and kept in g_noalloc_ptr.
static constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[kMaxDevices] = {};

// The current device's copy of the scratch array (looked up on the first
// call on that device); the stream must belong to the current device.
int noalloc_scratch(void *stream, void **out)
{
    int dev = -1;
    FRH_HIP(hipGetDevice(&dev), "frh_prefilter_noalloc: current device");
    FRH_REQ(dev >= 0 && dev < kMaxDevices, "frh_prefilter_noalloc: device %d out of range", dev);
    if (stream) {
        hipDevice_t sdev = -1;
        FRH_HIP(hipStreamGetDevice(as_stream(stream), &sdev), "frh_prefilter_noalloc: stream device");
        FRH_REQ((int)sdev == dev, "frh_prefilter_noalloc: the stream belongs to device %d, device %d is current",
                (int)sdev, dev);
    }
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        FRH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "frh_prefilter_noalloc: scratch address");
        FRH_REQ(p != nullptr, "frh_prefilter_noalloc: no device address for the scratch array");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return FRH_OK;
}
}  // namespace

extern "C" {

const char *frh_last_error(void) { return g_err; }

int frh_abi_version(void) { return 10; }

int frh_fill_resize_tables(int in_len, int out_len, int R, const float *g, int32_t *idx, float *wt)
{
    FRH_REQ(in_len > 0 && out_len > 0 && R >= 0 && g && idx && wt, "frh_fill_resize_tables: bad arguments");
    int P = 2 * R + 4;
    double scale = (double)out_len / (double)in_len;
    for (int i = 0; i < out_len; ++i) {
        double x = (i + 0.5) / scale - 0.5;
        int left = (int)std::floor(x - 2.0) - R;
        double ssum = 0.0;
        for (int p = 0; p < P; ++p) {
            int j = left + p;
            idx[(size_t)i * P + p] = reflect_index(j, in_len);
            double d = x - j, acc = 0.0;
            for (int u = -R; u <= R; ++u) acc += (double)g[u + R] * keys_cubic(d - u);
            wt[(size_t)i * P + p] = (float)acc;
            ssum += acc;
        }
        double inv = 1.0 / ssum;
        for (int p = 0; p < P; ++p) wt[(size_t)i * P + p] = (float)((double)wt[(size_t)i * P + p] * inv);
    }
    return FRH_OK;
}

int frh_warping_depth(double eta, int levels, int m, int n)
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

// ------------------------------------------------------------------ stages

// Scratch of the prefilter: the float64 planes between the vertical and the
// horizontal pass, plus (temporal kernel) the float64 frames after the pass
// along T.
size_t frh_prefilter_scratch_bytes(int T, int H, int W, int C, int rt)
{
    if (T <= 0 || H <= 0 || W <= 0 || C <= 0 || rt < 0) return 0;
    const size_t plane = sizeof(double) * (size_t)T * C * H * W;
    return rt > 0 ? 2 * plane : plane;
}

#define FRH_SCRATCH(ptr_, have, need, what)                                                        \
    FRH_REQ((ptr_) != nullptr && (have) >= (need), "%s: scratch of %zu bytes needed, %zu given", what, \
            (size_t)(need), (size_t)((ptr_) ? (have) : 0))

static int prefilter_impl(const void *raw, int f64, int T, int H, int W, int C, const double *lo,
                          const double *den, const double *wy, int ry, const double *wx, int rx,
                          void *out, int out_f64, int sem, double *tmp, void *stream)
{
    FRH_REQ(raw && out && lo && den && wy && wx, "frh_prefilter: null pointer");
    FRH_REQ(T > 0 && H > 0 && W > 0 && C > 0 && C <= 32, "frh_prefilter: bad shape T=%d H=%d W=%d C=%d", T, H, W, C);
    FRH_REQ(ry >= 0 && ry <= 32 && rx >= 0 && rx <= 32, "frh_prefilter: Gaussian radius must be in [0, 32]");
    hipStream_t st = as_stream(stream);
    hipError_t e = f64 ? launch_prefilter_v64((const double *)raw, T, H, W, C, lo, den, wy, ry, sem, tmp, st)
                       : launch_prefilter_v((const float *)raw, T, H, W, C, lo, den, wy, ry, sem, tmp, st);
    if (e == hipSuccess)
        e = out_f64 ? launch_prefilter_h64(tmp, T, H, W, C, wx, rx, sem, (double *)out, st)
                    : launch_prefilter_h(tmp, T, H, W, C, wx, rx, sem, (float *)out, st);
    if (e != hipSuccess) return fail(FRH_EHIP, "prefilter kernels: %s", hipGetErrorString(e));
    return FRH_OK;
}

int frh_prefilter_ex(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                     const double *den, const double *wy, int ry, const double *wx, int rx, void *out,
                     int out_f64, int semantics, void *scratch, size_t scratch_bytes, void *stream)
{
    FRH_REQ((semantics & ~(FRH_PRE_NORM32 | FRH_PRE_FILT32)) == 0, "frh_prefilter_ex: unknown semantics flags");
    FRH_SCRATCH(scratch, scratch_bytes, frh_prefilter_scratch_bytes(T, H, W, C, 0), "frh_prefilter_ex");
    return prefilter_impl(raw, raw_f64 != 0, T, H, W, C, lo, den, wy, ry, wx, rx, out, out_f64 != 0, semantics,
                          (double *)scratch, stream);
}

size_t frh_prefilter_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int frh_prefilter_noalloc(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                          const double *den, const double *wy, int ry, const double *wx, int rx, void *out,
                          int out_f64, int semantics, void *stream)
{
    FRH_REQ((semantics & ~(FRH_PRE_NORM32 | FRH_PRE_FILT32)) == 0,
            "frh_prefilter_noalloc: unknown semantics flags");
    FRH_REQ(T > 0 && H > 0 && W > 0 && C > 0, "frh_prefilter_noalloc: bad shape T=%d H=%d W=%d C=%d", T, H, W, C);
    const size_t need = frh_prefilter_scratch_bytes(T, H, W, C, 0);
    FRH_REQ(need <= kNoAllocScratchBytes,
            "frh_prefilter_noalloc: %zu bytes of scratch needed (T*C*H*W float64 values), the static scratch "
            "holds %zu; split the frames over several calls or use frh_prefilter_ex",
            need, kNoAllocScratchBytes);
    void *p = nullptr;
    const int rc = noalloc_scratch(stream, &p);
    if (rc) return rc;
    // the vertical pass writes every scratch value the horizontal pass reads:
    // nothing of an earlier call's contents is used
    return prefilter_impl(raw, raw_f64 != 0, T, H, W, C, lo, den, wy, ry, wx, rx, out, out_f64 != 0, semantics,
                          (double *)p, stream);
}

int frh_prefilter_3d(const void *raw, int raw_f64, int T, int H, int W, int C, const double *lo,
                     const double *den, const double *wt, int rt, const double *wy, int ry, const double *wx,
                     int rx, void *out, int out_f64, void *scratch, size_t scratch_bytes, void *stream)
{
    FRH_REQ(raw && out && lo && den && wt && wy && wx, "frh_prefilter_3d: null pointer");
    FRH_REQ(rt >= 0 && rt <= 32, "frh_prefilter_3d: temporal radius must be in [0, 32]");
    FRH_REQ(T > 0 && H > 0 && W > 0 && C > 0 && C <= 32 && H <= 65535 && T <= 65535,
            "frh_prefilter_3d: bad shape T=%d H=%d W=%d C=%d", T, H, W, C);
    FRH_SCRATCH(scratch, scratch_bytes, frh_prefilter_scratch_bytes(T, H, W, C, rt), "frh_prefilter_3d");
    double *tmp = (double *)scratch;
    if (rt == 0)  // the temporal kernel is [1]: an exact identity
        return prefilter_impl(raw, raw_f64 != 0, T, H, W, C, lo, den, wy, ry, wx, rx, out, out_f64 != 0, 0, tmp,
                              stream);
    hipStream_t st = as_stream(stream);
    double *tf = tmp + (size_t)T * C * H * W;
    hipError_t e = launch_prefilter_t(raw, raw_f64 != 0, T, H, W, C, lo, den, wt, rt, tf, st);
    if (e != hipSuccess) return fail(FRH_EHIP, "temporal prefilter kernel: %s", hipGetErrorString(e));
    // the frames are normalised now: x -> (x - 0) / 1 is exact
    std::vector<double> zero(C, 0.0), one(C, 1.0);
    return prefilter_impl(tf, 1, T, H, W, C, zero.data(), one.data(), wy, ry, wx, rx, out, out_f64 != 0, 0, tmp,
                          stream);
}

int frh_resize(const float *src, int N, int h, int w, float *dst, int oh, int ow, const int32_t *ix,
               const float *wx, const int32_t *iy, const float *wy, int P, float *tmp, void *stream)
{
    FRH_REQ(src && dst && ix && wx && iy && wy && tmp, "frh_resize: null pointer");
    FRH_REQ(N > 0 && h > 0 && w > 0 && oh > 0 && ow > 0 && P >= 4, "frh_resize: bad shape");
    FRH_REQ(oh <= 65535 && h <= 65535 && N <= 65535, "frh_resize: grid limits exceeded");
    hipError_t e = launch_resize(src, N, h, w, dst, oh, ow, ix, wx, iy, wy, P, tmp, as_stream(stream));
    FRH_HIP(e, "resize kernels");
    return FRH_OK;
}

int frh_warp_cubic(const float *f2, const float *f1, int f1_per_frame, int T, int C, int H, int W,
                   const float *u, const float *v, double h_col, double h_row, float *out, void *stream)
{
    FRH_REQ(f2 && f1 && u && v && out, "frh_warp_cubic: null pointer");
    FRH_REQ(T > 0 && C > 0 && H > 0 && W > 0 && H <= 65535 && T <= 65535, "frh_warp_cubic: bad shape");
    FRH_REQ(h_col > 0 && h_row > 0, "frh_warp_cubic: grid spacings must be positive");
    hipError_t e = launch_warp(f2, f1, f1_per_frame, T, C, H, W, u, v, (size_t)H * W, h_col, h_row, 0, out,
                               as_stream(stream));
    FRH_HIP(e, "warp kernel");
    return FRH_OK;
}

int frh_warp_cubic_ex(const void *f2, const void *f1, int img_f64, int acc_f64, int f1_per_frame, int T, int C,
                      int H, int W, const double *u, const double *v, float *out, void *stream)
{
    FRH_REQ(f2 && f1 && u && v && out, "frh_warp_cubic_ex: null pointer");
    FRH_REQ(T > 0 && C > 0 && H > 0 && W > 0 && H <= 65535 && T <= 65535, "frh_warp_cubic_ex: bad shape");
    FRH_REQ(!img_f64 || acc_f64, "frh_warp_cubic_ex: float64 images take cv2's 64F path (acc_f64 = 1)");
    hipError_t e = launch_warp_ex(f2, f1, img_f64 != 0, acc_f64 != 0, f1_per_frame, T, C, H, W, u, v,
                                  (size_t)H * W, out, as_stream(stream));
    FRH_HIP(e, "warp kernel");
    return FRH_OK;
}

int frh_warp_linear_ex(const void *f2, const void *f1, int img_f64, int acc_f64, int f1_per_frame, int T,
                       int C, int H, int W, const double *u, const double *v, float *out, void *stream)
{
    FRH_REQ(f2 && f1 && u && v && out, "frh_warp_linear_ex: null pointer");
    FRH_REQ(T > 0 && C > 0 && H > 0 && W > 0 && H <= 65535 && T <= 65535, "frh_warp_linear_ex: bad shape");
    FRH_REQ(!img_f64 || acc_f64, "frh_warp_linear_ex: float64 images take cv2's 64F path (acc_f64 = 1)");
    hipError_t e = launch_warp_linear_ex(f2, f1, img_f64 != 0, acc_f64 != 0, f1_per_frame, T, C, H, W, u, v,
                                         (size_t)H * W, out, as_stream(stream));
    FRH_HIP(e, "linear warp kernel");
    return FRH_OK;
}

int frh_warp_int16(int linear, int is_signed, const float *f2, const float *f1, int f1_per_frame, int T, int C,
                   int H, int W, const double *u, const double *v, float *out, void *stream)
{
    FRH_REQ(f2 && f1 && u && v && out, "frh_warp_int16: null pointer");
    FRH_REQ(T > 0 && C > 0 && H > 0 && W > 0 && H <= 65535 && T <= 65535, "frh_warp_int16: bad shape");
    const float lo = is_signed ? -32768.f : 0.f, hi = is_signed ? 32767.f : 65535.f;
    hipError_t e = launch_warp_saturate(linear != 0, f2, f1, f1_per_frame, T, C, H, W, u, v, (size_t)H * W, lo,
                                        hi, out, as_stream(stream));
    FRH_HIP(e, "16-bit warp kernel");
    return FRH_OK;
}

int frh_warp_u8(int linear, const float *f2, const float *f1, int f1_per_frame, int T, int C, int H, int W,
                const double *u, const double *v, float *out, void *stream)
{
    FRH_REQ(f2 && f1 && u && v && out, "frh_warp_u8: null pointer");
    FRH_REQ(T > 0 && C > 0 && H > 0 && W > 0 && H <= 65535 && T <= 65535, "frh_warp_u8: bad shape");
    FRH_HIP(launch_warp_u8(linear != 0, f2, f1, f1_per_frame, T, C, H, W, u, v, (size_t)H * W, out,
                           as_stream(stream)),
            "8-bit warp kernel");
    return FRH_OK;
}

int frh_motion_tensor(int kind, const float *f1, const float *f2, int h, int w, double hy, double hx,
                      double eps, int f1_f32, int f2_f32, double *J11, double *J22, double *J33,
                      double *J12, double *J13, double *J23, void *stream)
{
    FRH_REQ(f1 && f2 && J11 && J22 && J33 && J12 && J13 && J23, "frh_motion_tensor: null pointer");
    FRH_REQ(kind >= FRH_TENSOR_GC && kind <= FRH_TENSOR_CS, "frh_motion_tensor: unknown data term %d", kind);
    FRH_REQ(h > 0 && w > 0 && h + 2 <= 65535, "frh_motion_tensor: bad shape");
    FRH_REQ(hy > 0 && hx > 0, "frh_motion_tensor: grid spacings must be positive");
    FRH_REQ(kind != FRH_TENSOR_CS || eps > 0, "frh_motion_tensor: census eps must be positive");
    hipError_t e = launch_tensor_rowmajor(kind, f1_f32, f2_f32, f1, f2, h, w, hy, hx, eps, J11, J22, J33,
                                          J12, J13, J23, as_stream(stream));
    FRH_HIP(e, "tensor kernel");
    return FRH_OK;
}

size_t frh_resize_cubic_cv_scratch_bytes(int T, int h, int ow, int C)
{
    if (T <= 0 || h <= 0 || ow <= 0 || C <= 0) return 0;
    return sizeof(double) * (size_t)T * h * ow * C;
}

int frh_resize_cubic_cv(const double *src, int T, int h, int w, int C, double *dst, int oh, int ow,
                        void *scratch, size_t scratch_bytes, void *stream)
{
    FRH_REQ(src && dst && T > 0 && h > 0 && w > 0 && C > 0 && oh > 0 && ow > 0, "frh_resize_cubic_cv: bad arguments");
    FRH_REQ(T <= 65535 && oh <= 65535 && h <= 65535, "frh_resize_cubic_cv: grid limits exceeded");
    FRH_SCRATCH(scratch, scratch_bytes, frh_resize_cubic_cv_scratch_bytes(T, h, ow, C), "frh_resize_cubic_cv");
    FRH_HIP(launch_cv_resize(src, T, h, w, C, (double *)scratch, dst, oh, ow, as_stream(stream)),
            "cubic resize kernels");
    return FRH_OK;
}

int frh_mean_frames_f32(const float *w, int n, size_t len, float *out, void *stream)
{
    FRH_REQ(w && out && n > 0 && len > 0, "frh_mean_frames_f32: bad arguments");
    FRH_HIP(launch_mean_frames_f32(w, n, len, out, as_stream(stream)), "frame mean kernel");
    return FRH_OK;
}

size_t frh_median5_mirror_scratch_bytes(int N, int h, int w)
{
    if (N <= 0 || h <= 0 || w <= 0) return 0;
    return sizeof(double) * (size_t)N * h * w;
}

int frh_median5_mirror(double *f, int N, int h, int w, void *scratch, size_t scratch_bytes, void *stream)
{
    FRH_REQ(f && N > 0 && h > 0 && w > 0 && h <= 65535 && N <= 65535, "frh_median5_mirror: bad arguments");
    FRH_SCRATCH(scratch, scratch_bytes, frh_median5_mirror_scratch_bytes(N, h, w), "frh_median5_mirror");
    FRH_HIP(launch_median_plane(f, N, h, w, (double *)scratch, as_stream(stream)), "median kernel");
    return FRH_OK;
}

// Scratch of the stand-alone level solver: the skewed planes of J (6 K),
// the weights (K), cu / cv, the coefficients (5, or 9 with a_smooth != 1),
// du / dv and the border ring, plus psi_s on the padded grid when
// a_smooth != 1.
size_t frh_level_solver_scratch_bytes(int m, int n, int K, double a_smooth)
{
    if (m < 3 || n < 3 || K < 1) return 0;
    Skew s = make_skew(m - 2, n - 2);
    const bool nl = a_smooth != 1.0;
    const size_t ncoef = nl ? 9 : 5, nps = nl ? (size_t)m * n : 0;
    const size_t npl = (size_t)6 * K + K + 2 + ncoef + 2 + 2;
    return sizeof(double) * (npl * s.plane + nps);
}

int frh_level_solver(const double *J11, const double *J22, const double *J33, const double *J12,
                     const double *J13, const double *J23, const double *weight, const double *u,
                     const double *v, int m, int n, int K, double alpha_x, double alpha_y,
                     int iterations, int update_lag, const double *a_data, double a_smooth,
                     double hx, double hy, double gnc_beta, double *du, double *dv, void *scratch,
                     size_t scratch_bytes, void *stream)
{
    FRH_REQ(J11 && J22 && J33 && J12 && J13 && J23 && weight && u && v && a_data && du && dv,
            "frh_level_solver: null pointer");
    FRH_REQ(m >= 3 && n >= 3 && K >= 1 && K <= 8, "frh_level_solver: need m, n >= 3 and 1 <= K <= 8 (m=%d n=%d K=%d)", m, n, K);
    FRH_REQ(iterations >= 0 && update_lag >= 1, "frh_level_solver: iterations >= 0, update_lag >= 1");
    FRH_REQ(hx > 0 && hy > 0, "frh_level_solver: hx, hy must be positive");
    hipStream_t st = as_stream(stream);
    Skew s = make_skew(m - 2, n - 2);
    const double *Js[6] = {J11, J22, J33, J12, J13, J23};
    // a_smooth != 1: nine coefficients per cell and psi_s on the padded grid
    const bool nl = a_smooth != 1.0;
    const size_t ncoef = nl ? 9 : 5, nps = nl ? (size_t)m * n : 0;
    size_t npl = (size_t)6 * K + K + 2 + ncoef + 2 + 2;
    FRH_SCRATCH(scratch, scratch_bytes, frh_level_solver_scratch_bytes(m, n, K, a_smooth), "frh_level_solver");
    double *work = (double *)scratch;
    double *J = work, *Wsk = J + 6 * K * s.plane, *cuv = Wsk + K * s.plane, *coef = cuv + 2 * s.plane;
    double *duv = coef + ncoef * s.plane, *ring = duv + 2 * s.plane, *psis = work + npl * s.plane;
    double ax = alpha_x / (hx * hx), ay = alpha_y / (hy * hy);
    FRH_HIP(launch_fill_f64(work, npl * s.plane, 0.0, st), "level solver clear");
    if (nl) FRH_HIP(launch_fill_f64(psis, nps, 1.0, st), "psi_s init");
    std::vector<double> am1(K);
    for (int k = 0; k < K; ++k) am1[k] = a_data[k] - 1.0;
    hipError_t e = hipSuccess;
    for (int c = 0; c < 6 && e == hipSuccess; ++c)
        e = launch_padded_to_skew_f64(Js[c], m, n, K, s, J + (size_t)c * K * s.plane, st);
    if (e == hipSuccess) e = launch_padded_to_skew_f64(weight, m, n, K, s, Wsk, st);
    if (e == hipSuccess) e = launch_cuv_padded(u, v, m, n, ax, ay, s, cuv, st);
    const int pmax = nl ? sweep_pmax_nl(s.h) : sweep_pmax(s.h);
    std::vector<Pass> passes = plan_passes(iterations, update_lag, pmax);
    size_t fstride = npl * s.plane;
    if (passes.empty() && e == hipSuccess) e = launch_fill_f64(duv, 4 * s.plane, 0.0, st);
    const bool fused = !nl && fused_sweep_supported(K) && passes_fusable(passes);
    for (size_t k = 0; k < passes.size() && e == hipSuccess; ++k) {
        double *rg = k + 1 == passes.size() ? ring : nullptr;
        if (nl) {
            // every pass records the border cells' start values: the ring
            // that the next psi_s refresh sees
            const Pass &p = passes[k];
            if (p.assemble && !p.psi_one)
                e = launch_psi_smooth(nullptr, nullptr, 0, u, v, duv, ring, fstride, 1, s, k == 0 ? 1 : 0,
                                      a_smooth, gnc_beta, hx, hy, psis, 0, st);
            if (e == hipSuccess && p.assemble)
                e = launch_assemble_nl(J, Wsk, 0, nullptr, nullptr, 0, u, v, psis, 0, duv, coef, fstride, 1, K, s,
                                       ax, ay, am1.data(), a_data, p.psi_one ? 1 : 0, k == 0 ? 1 : 0, gnc_beta,
                                       st);
            if (e == hipSuccess) e = launch_sweep_nl(p.nsweeps, coef, duv, ring, fstride, 1, s, k == 0 ? 1 : 0, st);
            continue;
        }
        if (fused) {
            e = launch_sweep_fused(passes[k].nsweeps, J, Wsk, cuv, duv, rg, fstride, 1, K, s, ax, ay,
                                   am1.data(), a_data, passes[k].psi_one ? 1 : 0, k == 0 ? 1 : 0,
                                   gnc_beta, st);
            continue;
        }
        if (passes[k].assemble)
            e = launch_assemble(J, Wsk, cuv, duv, coef, fstride, 1, K, s, ax, ay, am1.data(), a_data,
                                passes[k].psi_one ? 1 : 0, k == 0 ? 1 : 0, gnc_beta, st);
        if (e == hipSuccess)
            e = launch_sweep(passes[k].nsweeps, coef, duv, rg, fstride, 1, s, ax, ay, k == 0 ? 1 : 0, st);
    }
    if (e == hipSuccess) e = launch_skew_to_padded_f64(duv, ring, s, du, dv, st);
    FRH_HIP(e, "level solver kernels");
    return FRH_OK;
}

}  // extern "C"

// ------------------------------------------------------------- whole pyramid

struct LevelDev {
    int h, w;
    double h_row, h_col, axw, ayw;
    int P_img, P_up;
    int32_t *img_ix = nullptr, *img_iy = nullptr, *up_ix = nullptr, *up_iy = nullptr;
    float *img_wx = nullptr, *img_wy = nullptr, *up_wx = nullptr, *up_wy = nullptr;
    Skew sk;
    size_t f1_off, W_off;
    std::vector<Pass> passes;
};

enum { TM_TOTAL, TM_SWEEP, TM_ASSEMBLE, TM_TENSOR, TM_RESIZE, TM_WARP, TM_MEDIAN, TM_N };

struct frh_solver {
    int H, W, C, nl, iterations, update_lag, max_batch;
    int const_assumption = FRH_TENSOR_GC;
    double a_smooth;
    std::vector<double> a_data, am1;
    std::vector<LevelDev> lv;
    float *f1lev = nullptr, *uvtop = nullptr, *tmp = nullptr, *f2l = nullptr, *f2w = nullptr;
    float *uvl = nullptr, *uvn = nullptr;
    float *Wlev = nullptr;  // weight levels, skewed float32 (float32 values)
    double *work = nullptr;
    double *psis = nullptr;       // a_smooth != 1: psi_s per frame on the padded level grid
    int Wsum = 0;                 // sum of level widths (multi-level horizontal pass)
    std::vector<int> xoff;        // per level: column offset in the staged rows
    int32_t *hoff = nullptr, *hix = nullptr;
    float *hwx = nullptr, *pyr = nullptr;
    double *wlow = nullptr, *wtmp = nullptr;  // min_level > 0: coarse flow, resize scratch
    float *pyr1 = nullptr, *f1l = nullptr;    // per-frame references: pyramid rows, level images
    bool have_fixed = false;
    size_t npl = 0;
    bool have_ref = false;
    bool profile = false;
    struct Mark {
        int cat, level;
        hipEvent_t a, b;
    };
    int cur_level = -1;                 // level being launched (profile marks)
    std::vector<double> level_ms;       // sweep + assemble time per level of the last profiled run
    std::vector<Mark> marks;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    double tm[TM_N] = {0};
    int counts[TM_N] = {0};
    // graph replay of the level loop (frh_solver_set_graph): captured once per
    // (T, output type, per-frame references) on a solver-owned stream, with
    // the caller's frames and output staged through solver-owned buffers
    bool use_graph = false;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t gexec = nullptr;
    int g_T = 0, g_out_f64 = 0, g_pairs = 0;
    float *stage_in = nullptr, *stage_fixed = nullptr;
    void *stage_out = nullptr;
    // small batches: a pass's sweeps as a chain of work-groups (one per
    // sweep and frame) when they fit on the device at once; progress flags,
    // the failure word (host memory mapped to the device: set by a stage
    // that gave up waiting, checked at the next run) and the work-group
    // slots of the chain kernel
    int *chain_flags = nullptr;
    int *chain_err = nullptr, *chain_err_dev = nullptr;
    int chain_slots = 0;
    bool use_chains = true;
};

static void free_solver(frh_solver *s)
{
    if (!s) return;
    for (auto &L : s->lv) {
        (void)hipFree(L.img_ix); (void)hipFree(L.img_iy); (void)hipFree(L.img_wx); (void)hipFree(L.img_wy);
        (void)hipFree(L.up_ix); (void)hipFree(L.up_iy); (void)hipFree(L.up_wx); (void)hipFree(L.up_wy);
    }
    (void)hipFree(s->f1lev); (void)hipFree(s->uvtop); (void)hipFree(s->tmp); (void)hipFree(s->f2l); (void)hipFree(s->f2w);
    (void)hipFree(s->uvl); (void)hipFree(s->uvn); (void)hipFree(s->Wlev); (void)hipFree(s->work); (void)hipFree(s->psis);
    (void)hipFree(s->hoff); (void)hipFree(s->hix); (void)hipFree(s->hwx); (void)hipFree(s->pyr);
    (void)hipFree(s->pyr1); (void)hipFree(s->f1l);
    (void)hipFree(s->wlow); (void)hipFree(s->wtmp);
    for (auto e : s->pool) (void)hipEventDestroy(e);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->cap_stream) (void)hipStreamDestroy(s->cap_stream);
    (void)hipFree(s->stage_in); (void)hipFree(s->stage_fixed); (void)hipFree(s->stage_out);
    (void)hipFree(s->chain_flags);
    if (s->chain_err) (void)hipHostFree(s->chain_err);
    delete s;
}

// Event bracketing of one launch group (only when profiling is enabled).
static hipEvent_t next_event(frh_solver *s)
{
    if (s->pool_used == s->pool.size()) {
        hipEvent_t e;
        (void)hipEventCreate(&e);
        s->pool.push_back(e);
    }
    return s->pool[s->pool_used++];
}

struct Scope {
    frh_solver *s;
    int cat;
    hipStream_t st;
    hipEvent_t a{};
    Scope(frh_solver *s_, int cat_, hipStream_t st_) : s(s_), cat(cat_), st(st_)
    {
        if (s->profile) {
            a = next_event(s);
            (void)hipEventRecord(a, st);
        }
    }
    ~Scope()
    {
        if (s->profile) {
            hipEvent_t b = next_event(s);
            (void)hipEventRecord(b, st);
            s->marks.push_back({cat, s->cur_level, a, b});
        }
    }
};

// The level loop for T moving frames against the shared reference levels
// (fixed == nullptr) or against per-frame references fixed (T, C, H, W).
static int solver_run(frh_solver *s, const float *fixed, const float *moving, int T, void *w_out,
                      int out_f64, hipStream_t st);
static int solver_dispatch(frh_solver *s, const float *fixed, const float *moving, int T, void *w_out,
                           int out_f64, hipStream_t st);

extern "C" {

int frh_solver_create(frh_solver **out, const frh_solver_desc *d)
{
    FRH_REQ(out && d && d->levels && d->a_data, "frh_solver_create: null pointer");
    FRH_REQ(d->H > 0 && d->W > 0 && d->C >= 1 && d->C <= 8, "frh_solver_create: bad geometry");
    FRH_REQ(d->n_levels >= 1 && d->max_batch >= 1 && d->max_batch <= 65535, "frh_solver_create: bad level count or batch");
    FRH_REQ(d->iterations >= 0 && d->update_lag >= 1, "frh_solver_create: iterations >= 0, update_lag >= 1");
    FRH_REQ(d->const_assumption >= FRH_TENSOR_GC && d->const_assumption <= FRH_TENSOR_CS,
            "frh_solver_create: unknown data term %d", d->const_assumption);
    frh_solver *s = new frh_solver();
    s->H = d->H; s->W = d->W; s->C = d->C; s->nl = d->n_levels; s->max_batch = d->max_batch;
    s->iterations = d->iterations; s->update_lag = d->update_lag; s->a_smooth = d->a_smooth;
    s->const_assumption = d->const_assumption;
    s->a_data.assign(d->a_data, d->a_data + d->C);
    for (double a : s->a_data) s->am1.push_back(a - 1.0);
    size_t f1_total = 0, W_total = 0, plane_max = 0;
    int hmax = 0, wmax = 0;
    int rc = FRH_OK;
    for (int i = 0; i < d->n_levels && rc == FRH_OK; ++i) {
        const frh_level_desc &L = d->levels[i];
        if (L.h < 1 || L.w < 1 || L.h > 65535 || L.P_img < 4 || (i > 0 && L.P_up < 4)) {
            rc = fail(FRH_EINVAL, "frh_solver_create: level %d invalid (h=%d w=%d)", i, L.h, L.w);
            break;
        }
        if (i > 0 && (L.h < d->levels[i - 1].h || L.w < d->levels[i - 1].w)) {
            rc = fail(FRH_EINVAL, "frh_solver_create: levels must go from coarse to fine");
            break;
        }
        LevelDev D;
        D.h = L.h; D.w = L.w; D.h_row = L.h_row; D.h_col = L.h_col;
        D.axw = L.alpha_x / (L.h_col * L.h_col);
        D.ayw = L.alpha_y / (L.h_row * L.h_row);
        D.P_img = L.P_img; D.P_up = i > 0 ? L.P_up : 0;
        D.sk = make_skew(L.h, L.w);
        D.f1_off = f1_total;
        D.W_off = W_total;
        D.passes = plan_passes(d->iterations, d->update_lag,
                               d->a_smooth != 1.0 ? sweep_pmax_nl(L.h) : sweep_pmax(L.h));
        f1_total += (size_t)d->C * L.h * L.w;
        W_total += (size_t)d->C * D.sk.plane;
        if (D.sk.plane > plane_max) plane_max = D.sk.plane;
        if (L.h > hmax) hmax = L.h;
        if (L.w > wmax) wmax = L.w;
        rc = upload(&D.img_ix, L.img_ix, (size_t)L.w * L.P_img, "tables");
        if (!rc) rc = upload(&D.img_wx, L.img_wx, (size_t)L.w * L.P_img, "tables");
        if (!rc) rc = upload(&D.img_iy, L.img_iy, (size_t)L.h * L.P_img, "tables");
        if (!rc) rc = upload(&D.img_wy, L.img_wy, (size_t)L.h * L.P_img, "tables");
        if (!rc && i > 0) {
            rc = upload(&D.up_ix, L.up_ix, (size_t)L.w * L.P_up, "tables");
            if (!rc) rc = upload(&D.up_wx, L.up_wx, (size_t)L.w * L.P_up, "tables");
            if (!rc) rc = upload(&D.up_iy, L.up_iy, (size_t)L.h * L.P_up, "tables");
            if (!rc) rc = upload(&D.up_wy, L.up_wy, (size_t)L.h * L.P_up, "tables");
        }
        s->lv.push_back(D);
    }
    if (!rc) {
        std::vector<int32_t> hoff(1, 0), hix;
        std::vector<float> hwx;
        for (int i = 0; i < d->n_levels; ++i) {
            const frh_level_desc &L = d->levels[i];
            s->xoff.push_back(s->Wsum);
            s->Wsum += L.w;
            for (int x = 0; x < L.w; ++x) {
                for (int p = 0; p < L.P_img; ++p) {
                    hix.push_back(L.img_ix[(size_t)x * L.P_img + p]);
                    hwx.push_back(L.img_wx[(size_t)x * L.P_img + p]);
                }
                hoff.push_back((int32_t)hix.size());
            }
        }
        rc = upload(&s->hoff, hoff.data(), hoff.size(), "pyramid tables");
        if (!rc) rc = upload(&s->hix, hix.data(), hix.size(), "pyramid tables");
        if (!rc) rc = upload(&s->hwx, hwx.data(), hwx.size(), "pyramid tables");
    }
    size_t T = (size_t)d->max_batch, C = d->C, HW = (size_t)d->H * d->W;
    const bool nl = d->a_smooth != 1.0;
    // J, cu/cv, coefficients (9 with a_smooth != 1), du/dv, border ring
    s->npl = 6 * C + 2 + (nl ? 9 : 5) + 2 + (nl ? 2 : 0);
    size_t cmax = C > 2 ? C : 2;
    if (!rc) rc = dev_alloc(&s->f1lev, f1_total, "reference levels");
    if (!rc) rc = dev_alloc(&s->Wlev, W_total, "weight levels");
    if (!rc) rc = dev_alloc(&s->uvtop, (size_t)2 * s->lv[0].h * s->lv[0].w, "initial flow");
    if (!rc) rc = dev_alloc(&s->tmp, T * cmax * HW, "resize scratch");
    if (!rc) rc = dev_alloc(&s->f2l, T * C * (size_t)hmax * wmax, "level images");
    if (!rc) rc = dev_alloc(&s->f2w, T * C * (size_t)hmax * wmax, "warped images");
    if (!rc) rc = dev_alloc(&s->uvl, T * 2 * (size_t)hmax * wmax, "level flow");
    if (!rc) rc = dev_alloc(&s->uvn, T * 2 * (size_t)hmax * wmax, "next flow");
    if (!rc) rc = dev_alloc(&s->work, T * s->npl * plane_max, "solver workspace");
    if (!rc && nl) rc = dev_alloc(&s->psis, T * (size_t)(hmax + 2) * (wmax + 2), "smoothness weights");
    if (!rc) rc = dev_alloc(&s->pyr, T * C * (size_t)d->H * s->Wsum, "pyramid rows");
    if (!rc && (hmax != d->H || wmax != d->W)) {
        rc = dev_alloc(&s->wlow, T * 2 * (size_t)hmax * wmax, "coarse flow");
        if (!rc) rc = dev_alloc(&s->wtmp, T * 2 * (size_t)hmax * d->W, "final resize scratch");
    }
    if (!rc) {
        s->chain_slots = sweep_chain_slots();
        if (s->chain_slots >= 2) {
            // at most chain_slots frames x stages, frames padded to a multiple of 8
            const size_t nflags = std::max((size_t)8 * s->chain_slots, (size_t)5 * chain_tpad(d->max_batch));
            rc = dev_alloc(&s->chain_flags, nflags * kChainFlagStride, "chain flags");
            hipError_t e = rc ? hipSuccess : hipHostMalloc((void **)&s->chain_err, 256, hipHostMallocMapped);
            if (!rc && e == hipSuccess) e = hipHostGetDevicePointer((void **)&s->chain_err_dev, s->chain_err, 0);
            if (!rc && e != hipSuccess) rc = fail(FRH_EHIP, "chain failure word: %s", hipGetErrorString(e));
            if (!rc) *s->chain_err = 0;
        }
    }
    if (rc) {
        free_solver(s);
        return rc;
    }
    *out = s;
    return FRH_OK;
}

void frh_solver_destroy(frh_solver *s) { free_solver(s); }

int frh_solver_set_reference(frh_solver *s, const float *fixed, const float *weight, const float *uv,
                             void *stream)
{
    FRH_REQ(s && weight, "frh_solver_set_reference: null pointer");
    hipStream_t st = as_stream(stream);
    int C = s->C;
    if (fixed) {
        FRH_HIP(launch_pyr_h(fixed, C, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr, st),
                "reference pyramid");
        for (int i = 0; i < s->nl; ++i) {
            LevelDev &L = s->lv[i];
            FRH_HIP(launch_pyr_v(s->pyr, C, s->H, s->Wsum, s->xoff[i], s->f1lev + L.f1_off, L.h, L.w,
                                 L.img_iy, L.img_wy, L.P_img, st), "reference pyramid");
        }
    }
    FRH_HIP(launch_pyr_h(weight, C, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr, st),
            "weight pyramid");
    for (int i = 0; i < s->nl; ++i) {
        LevelDev &L = s->lv[i];
        float *wl = s->f2l;  // scratch for the resized weight planes
        FRH_HIP(launch_pyr_v(s->pyr, C, s->H, s->Wsum, s->xoff[i], wl, L.h, L.w, L.img_iy, L.img_wy,
                             L.P_img, st), "weight pyramid");
        FRH_HIP(launch_to_skew_f32f(wl, C, L.sk, s->Wlev + L.W_off, st), "weight skew");
    }
    LevelDev &T0 = s->lv[0];
    if (uv) {
        FRH_HIP(launch_resize(uv, 2, s->H, s->W, s->uvtop, T0.h, T0.w, T0.img_ix, T0.img_wx, T0.img_iy,
                              T0.img_wy, T0.P_img, s->tmp, st), "initial flow resize");
    } else {
        FRH_HIP(hipMemsetAsync(s->uvtop, 0, sizeof(float) * 2 * T0.h * T0.w, st), "initial flow clear");
    }
    s->have_ref = true;
    s->have_fixed = fixed != nullptr;
    return FRH_OK;
}

int frh_solver_run(frh_solver *s, const float *moving, int T, void *w_out, int out_f64, void *stream)
{
    FRH_REQ(s && moving && w_out, "frh_solver_run: null pointer");
    FRH_REQ(s->have_ref && s->have_fixed, "frh_solver_run: frh_solver_set_reference has not set a reference");
    FRH_REQ(T >= 1 && T <= s->max_batch, "frh_solver_run: T=%d outside [1, %d]", T, s->max_batch);
    return solver_dispatch(s, nullptr, moving, T, w_out, out_f64, as_stream(stream));
}

int frh_solver_run_pairs(frh_solver *s, const float *fixed, const float *moving, int T, void *w_out,
                         int out_f64, void *stream)
{
    FRH_REQ(s && fixed && moving && w_out, "frh_solver_run_pairs: null pointer");
    FRH_REQ(s->have_ref, "frh_solver_run_pairs: frh_solver_set_reference has not been called");
    FRH_REQ(T >= 1 && T <= s->max_batch, "frh_solver_run_pairs: T=%d outside [1, %d]", T, s->max_batch);
    if (!s->pyr1) {
        size_t T_ = (size_t)s->max_batch, C = s->C;
        int rc = dev_alloc(&s->pyr1, T_ * C * (size_t)s->H * s->Wsum, "reference pyramid rows");
        if (!rc) rc = dev_alloc(&s->f1l, T_ * C * (size_t)s->lv.back().h * s->lv.back().w, "reference levels");
        if (rc) return rc;
    }
    return solver_dispatch(s, fixed, moving, T, w_out, out_f64, as_stream(stream));
}

}  // extern "C"

// Rows of a band in band mode (levels of more than 1024 rows).
constexpr int kBandRowsHost = 512;

// Levels with fewer anti-diagonals run their passes pipelined: the lag
// between chain stages (a few publication intervals) would be a large part
// of the pass.
constexpr int kChainMinDiagonals = 256;

static int solver_run(frh_solver *s, const float *fixed, const float *moving, int T, void *w_out,
                      int out_f64, hipStream_t st)
{
    if (s->chain_err && *(volatile int *)s->chain_err)
        return fail(FRH_EHIP, "a chained sweep stage of an earlier run gave up waiting for its predecessor "
                              "(that run's flows are invalid)");
    const int C = s->C;
    s->marks.clear();
    s->pool_used = 0;
    s->cur_level = -1;
    hipEvent_t t0{}, t1{};
    if (s->profile) {
        t0 = next_event(s);
        (void)hipEventRecord(t0, st);
    }
    {
        Scope sc(s, TM_RESIZE, st);
        FRH_HIP(launch_pyr_h(moving, T * C, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr, st),
                "moving pyramid rows");
        if (fixed)
            FRH_HIP(launch_pyr_h(fixed, T * C, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr1, st),
                    "reference pyramid rows");
    }
    for (int li = 0; li < s->nl; ++li) {
        LevelDev &L = s->lv[li];
        s->cur_level = li;
        const int h = L.h, w = L.w;
        const size_t hw = (size_t)h * w;
        const Skew &sk = L.sk;
        const size_t fs = s->npl * sk.plane;
        const bool nl = s->a_smooth != 1.0;
        double *J = s->work, *cuv = J + 6 * C * sk.plane, *coef = cuv + 2 * sk.plane;
        double *duv = coef + (nl ? 9 : 5) * sk.plane, *ring = duv + 2 * sk.plane;
        {
            Scope sc(s, TM_RESIZE, st);
            FRH_HIP(launch_pyr_v(s->pyr, T * C, s->H, s->Wsum, s->xoff[li], s->f2l, h, w, L.img_iy,
                                 L.img_wy, L.P_img, st), "level images");
            if (fixed)
                FRH_HIP(launch_pyr_v(s->pyr1, T * C, s->H, s->Wsum, s->xoff[li], s->f1l, h, w, L.img_iy,
                                     L.img_wy, L.P_img, st), "reference level images");
        }
        // reference level images: shared, or one per frame (T, C, h, w)
        const float *f1 = fixed ? s->f1l : s->f1lev + L.f1_off;
        const size_t f1_fs = fixed ? (size_t)C * hw : 0;
        const float *f2src = s->f2l;
        if (li == 0) {
            FRH_HIP(launch_broadcast_f32(s->uvtop, 2 * hw, s->uvl, 2 * hw, T, st), "initial flow");
        } else {
            const LevelDev &Lp = s->lv[li - 1];
            {
                Scope sc(s, TM_RESIZE, st);
                if (L.P_up == 4)
                    FRH_HIP(launch_resize_fused4(s->uvn, T * 2, Lp.h, Lp.w, s->uvl, h, w, L.up_ix, L.up_wx, L.up_iy,
                                                 L.up_wy, st), "flow upsample");
                else
                    FRH_HIP(launch_resize(s->uvn, T * 2, Lp.h, Lp.w, s->uvl, h, w, L.up_ix, L.up_wx, L.up_iy,
                                          L.up_wy, L.P_up, s->tmp, st), "flow upsample");
            }
            {
                Scope sc(s, TM_WARP, st);
                // the reference's level images are float64: cv2.remap's 64F path
                FRH_HIP(launch_warp(s->f2l, f1, fixed ? 1 : 0, T, C, h, w, s->uvl, s->uvl + hw, 2 * hw,
                                    L.h_col, L.h_row, 1, s->f2w, st), "warp");
            }
            f2src = s->f2w;
        }
        {
            Scope sc(s, TM_TENSOR, st);
            // the top level's "warped" image is the resized level itself
            // (float64 data in the reference); below it the warp's float32
            FRH_HIP(launch_tensor(s->const_assumption, li == 0 ? 0 : 1, f1, f1_fs, f2src, T, C, h, w,
                                  L.h_row, L.h_col, kCensusEps, s->uvl, s->uvl + hw, 2 * hw, L.axw, L.ayw,
                                  J, cuv, fs, sk, st),
                    "motion tensor");
        }
        // The separate assemble kernel is faster than forming coefficients
        // inside the wavefront (measured on MI355X), so the pyramid solver
        // uses it; frh_level_solver keeps the fused form.
        const size_t ps_stride = (size_t)(h + 2) * (w + 2);
        if (nl && !L.passes.empty() && L.passes[0].psi_one)
            FRH_HIP(launch_fill_f64(s->psis, (size_t)T * ps_stride, 1.0, st), "psi_s init");
        // levels of more than 512 rows: a lag block's passes as chains of
        // work-groups (sweeps x 512-row bands) when two sweeps' bands of all
        // frames fit one work-group per compute unit (1024 x 1024 x 2 frames:
        // 128 against 93 frames/s at 32 frames, 165 against 161 at 64; at
        // 128 frames, one sweep per chain launch, 199 against 253)
        const int nbands = (h + kBandRowsHost - 1) / kBandRowsHost;
        const bool chains = s->chain_err && s->use_chains;
        const bool band_chain = chains && sk.hs > 512 && 2 * T * nbands <= s->chain_slots;
        for (size_t k = 0; k < L.passes.size(); ++k) {
            const Pass &p = L.passes[k];
            if (band_chain && p.assemble) {
                if (nl) {
                    Scope sc(s, TM_ASSEMBLE, st);
                    if (!p.psi_one)
                        FRH_HIP(launch_psi_smooth(s->uvl, s->uvl + hw, 2 * hw, nullptr, nullptr, duv, ring, fs, T, sk,
                                                  k == 0 ? 1 : 0, s->a_smooth, -1.0, L.h_col, L.h_row, s->psis,
                                                  ps_stride, st), "smoothness weights");
                    FRH_HIP(launch_assemble_nl(J, s->Wlev + L.W_off, 1, s->uvl, s->uvl + hw, 2 * hw, nullptr, nullptr,
                                               s->psis, ps_stride, duv, coef, fs, T, C, sk, L.axw, L.ayw,
                                               s->am1.data(), s->a_data.data(), p.psi_one ? 1 : 0, k == 0 ? 1 : 0,
                                               -1.0, st), "assemble (nonlinear smoothness)");
                } else {
                    Scope sc(s, TM_ASSEMBLE, st);
                    FRH_HIP(launch_assemble_w32(J, s->Wlev + L.W_off, cuv, duv, coef, fs, T, C, sk, L.axw, L.ayw,
                                                s->am1.data(), s->a_data.data(), p.psi_one ? 1 : 0, k == 0 ? 1 : 0,
                                                -1.0, st), "assemble");
                }
                size_t m = 1;
                while (k + m < L.passes.size() && !L.passes[k + m].assemble) ++m;
                int left = 0;
                for (size_t i = 0; i < m; ++i) left += L.passes[k + i].nsweeps;
                const int per = s->chain_slots / (T * nbands);  // sweeps per launch
                int first = k == 0 ? 1 : 0;
                Scope sc(s, TM_SWEEP, st);
                while (left > 0) {
                    const int n = left < per ? left : per;
                    FRH_HIP(launch_sweep_band_chain(n, nl ? 1 : 0, coef, duv, nl ? ring : nullptr, fs, T, sk, L.axw,
                                                    L.ayw, first, s->chain_flags, s->chain_err_dev, st),
                            "sor sweep (band chain)");
                    first = 0;
                    left -= n;
                }
                k += m - 1;
                continue;
            }
            if (nl) {
                // a_smooth != 1: psi_s refresh, nine-coefficient assembly and
                // sweeps that record the border ring for the next refresh
                if (p.assemble) {
                    Scope sc(s, TM_ASSEMBLE, st);
                    if (!p.psi_one)
                        FRH_HIP(launch_psi_smooth(s->uvl, s->uvl + hw, 2 * hw, nullptr, nullptr, duv, ring, fs, T, sk,
                                                  k == 0 ? 1 : 0, s->a_smooth, -1.0, L.h_col, L.h_row, s->psis,
                                                  ps_stride, st), "smoothness weights");
                    FRH_HIP(launch_assemble_nl(J, s->Wlev + L.W_off, 1, s->uvl, s->uvl + hw, 2 * hw, nullptr, nullptr,
                                               s->psis, ps_stride, duv, coef, fs, T, C, sk, L.axw, L.ayw,
                                               s->am1.data(), s->a_data.data(), p.psi_one ? 1 : 0, k == 0 ? 1 : 0,
                                               -1.0, st), "assemble (nonlinear smoothness)");
                }
                Scope sc(s, TM_SWEEP, st);
                FRH_HIP(launch_sweep_nl(p.nsweeps, coef, duv, ring, fs, T, sk, k == 0 ? 1 : 0, st),
                        "sor sweep (nonlinear smoothness)");
                continue;
            }
            if (p.assemble) {
                Scope sc(s, TM_ASSEMBLE, st);
                FRH_HIP(launch_assemble_w32(J, s->Wlev + L.W_off, cuv, duv, coef, fs, T, C, sk, L.axw, L.ayw,
                                        s->am1.data(), s->a_data.data(), p.psi_one ? 1 : 0, k == 0 ? 1 : 0,
                                        -1.0, st), "assemble");
            }
            Scope sc(s, TM_SWEEP, st);
            // the fewest sweeps per chain stage (1 to 3, the last stage at
            // most one sweep short) whose stages fill at most half the
            // compute units, else no chain (512 x 512 x 2, frames/s, pipelined
            // pass / stages of 1 / 2 / 3 sweeps: 16 frames 264 / 328 / 315 /
            // -, 32 frames 486 / 519 / 560 / 533, 64 frames 840 / 826 / 847 /
            // 884, 128 frames 1292 / - / - / 1217)
            int per_stage = 0;
            for (int q = 1; q <= 3 && !per_stage; ++q) {
                const int stages = (p.nsweeps + q - 1) / q;
                if ((q == 1 || p.nsweeps % q == 0 || p.nsweeps % q == q - 1) && 2 * T * stages <= s->chain_slots)
                    per_stage = q;
            }
            if (chains && per_stage && sk.hs <= 512 && sk.h + sk.w - 1 >= kChainMinDiagonals && p.nsweeps >= 2)
                FRH_HIP(launch_sweep_chain(p.nsweeps, per_stage, coef, duv, nullptr, fs, T, sk, L.axw, L.ayw,
                                           k == 0 ? 1 : 0, s->chain_flags, s->chain_err_dev, st),
                        "sor sweep (chain)");
            else
                FRH_HIP(launch_sweep(p.nsweeps, coef, duv, nullptr, fs, T, sk, L.axw, L.ayw, k == 0 ? 1 : 0, st),
                        "sor sweep");
        }
        {
            Scope sc(s, TM_MEDIAN, st);
            const int med = (h < w ? h : w) > 5;
            if (li + 1 == s->nl && s->wlow) {
                // min_level > 0: float64 flow at the finest computed level, then
                // cv2.resize INTER_CUBIC to full resolution
                FRH_HIP(launch_median_add(duv, fs, sk, T, med, s->uvl, s->uvl + hw, 2 * hw, nullptr, s->wlow,
                                          1, st), "median/add (coarse output)");
                if (out_f64) {
                    FRH_HIP(launch_cv_resize(s->wlow, T, h, w, 2, s->wtmp, (double *)w_out, s->H, s->W, st),
                            "final cubic resize");
                } else {
                    return fail(FRH_EUNSUP, "min_level > 0 needs float64 output");
                }
            } else if (li + 1 == s->nl) {
                FRH_HIP(launch_median_add(duv, fs, sk, T, med, s->uvl, s->uvl + hw, 2 * hw, (float *)w_out,
                                          out_f64 ? (double *)w_out : nullptr, 1, st), "median/add (output)");
            } else {
                FRH_HIP(launch_median_add(duv, fs, sk, T, med, s->uvl, s->uvl + hw, 2 * hw, s->uvn, nullptr,
                                          0, st), "median/add");
            }
        }
    }
    if (s->profile) {
        t1 = next_event(s);
        (void)hipEventRecord(t1, st);
        FRH_HIP(hipStreamSynchronize(st), "profile sync");
        for (int k = 0; k < TM_N; ++k) { s->tm[k] = 0; s->counts[k] = 0; }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, t0, t1);
        s->tm[TM_TOTAL] = ms;
        s->counts[TM_TOTAL] = 1;
        s->level_ms.assign(s->nl, 0.0);
        for (auto &m : s->marks) {
            (void)hipEventElapsedTime(&ms, m.a, m.b);
            s->tm[m.cat] += ms;
            s->counts[m.cat] += 1;
            if ((m.cat == TM_SWEEP || m.cat == TM_ASSEMBLE) && m.level >= 0 && m.level < s->nl)
                s->level_ms[m.level] += ms;
        }
    }
    return FRH_OK;
}

// Eager launches, or a replay of the captured level loop.  The graph's
// kernels read and write the staging buffers, so one capture serves every
// call with the same (T, output type, per-frame references); the level
// images, reference and weights it uses live in the solver and keep their
// addresses across frh_solver_set_reference calls.
static int solver_dispatch(frh_solver *s, const float *fixed, const float *moving, int T, void *w_out,
                           int out_f64, hipStream_t st)
{
    if (!s->use_graph || s->profile) return solver_run(s, fixed, moving, T, w_out, out_f64, st);
    const size_t in_elems = (size_t)s->max_batch * s->C * s->H * s->W;
    const size_t n_in = (size_t)T * s->C * s->H * s->W;
    const size_t out_bytes = (size_t)T * s->H * s->W * 2 * (out_f64 ? sizeof(double) : sizeof(float));
    if (!s->stage_in) {
        int rc = dev_alloc(&s->stage_in, in_elems, "graph input staging");
        if (!rc) rc = dev_alloc((double **)&s->stage_out, (size_t)s->max_batch * s->H * s->W * 2,
                                "graph output staging");
        if (rc) return rc;
    }
    if (fixed && !s->stage_fixed) {
        int rc = dev_alloc(&s->stage_fixed, in_elems, "graph reference staging");
        if (rc) return rc;
    }
    const int pairs = fixed != nullptr;
    if (!s->gexec || s->g_T != T || s->g_out_f64 != out_f64 || s->g_pairs != pairs) {
        if (s->gexec) {
            (void)hipGraphExecDestroy(s->gexec);
            s->gexec = nullptr;
        }
        if (!s->cap_stream) FRH_HIP(hipStreamCreateWithFlags(&s->cap_stream, hipStreamNonBlocking), "capture stream");
        FRH_HIP(hipStreamBeginCapture(s->cap_stream, hipStreamCaptureModeThreadLocal), "begin capture");
        const int rc = solver_run(s, pairs ? s->stage_fixed : nullptr, s->stage_in, T, s->stage_out, out_f64,
                                  s->cap_stream);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->cap_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        FRH_HIP(e, "end capture");
        const hipError_t ei = hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FRH_HIP(ei, "graph instantiate");
        s->g_T = T;
        s->g_out_f64 = out_f64;
        s->g_pairs = pairs;
    }
    FRH_HIP(hipMemcpyAsync(s->stage_in, moving, sizeof(float) * n_in, hipMemcpyDeviceToDevice, st), "stage frames");
    if (pairs)
        FRH_HIP(hipMemcpyAsync(s->stage_fixed, fixed, sizeof(float) * n_in, hipMemcpyDeviceToDevice, st),
                "stage references");
    FRH_HIP(hipGraphLaunch(s->gexec, st), "graph launch");
    FRH_HIP(hipMemcpyAsync(w_out, s->stage_out, out_bytes, hipMemcpyDeviceToDevice, st), "unstage flows");
    return FRH_OK;
}

extern "C" {

int frh_solver_set_graph(frh_solver *s, int enable)
{
    FRH_REQ(s, "frh_solver_set_graph: null solver");
    s->use_graph = enable != 0;
    return FRH_OK;
}

int frh_solver_set_chains(frh_solver *s, int enable)
{
    FRH_REQ(s, "frh_solver_set_chains: null solver");
    if (s->use_chains != (enable != 0) && s->gexec) {  // the captured level loop holds the old launches
        (void)hipGraphExecDestroy(s->gexec);
        s->gexec = nullptr;
    }
    s->use_chains = enable != 0;
    return FRH_OK;
}

int frh_solver_profile(frh_solver *s, int enable)
{
    FRH_REQ(s, "frh_solver_profile: null solver");
    s->profile = enable != 0;
    return FRH_OK;
}

int frh_solver_timings(frh_solver *s, double *ms, int *counts, int n)
{
    FRH_REQ(s && ms && counts, "frh_solver_timings: null pointer");
    int k = 0;
    for (; k < n && k < TM_N; ++k) {
        ms[k] = s->tm[k];
        counts[k] = s->counts[k];
    }
    return k;
}

int frh_solver_level_info(frh_solver *s, int level, int *h, int *w, int *pass_limit, int *pass_limit_max,
                          double *last_ms)
{
    FRH_REQ(s && level >= 0 && level < s->nl, "frh_solver_level_info: bad level");
    const LevelDev &L = s->lv[level];
    int lim = 0;
    for (const Pass &p : L.passes) lim = p.nsweeps > lim ? p.nsweeps : lim;
    if (h) *h = L.h;
    if (w) *w = L.w;
    if (pass_limit) *pass_limit = lim;
    if (pass_limit_max) *pass_limit_max = sweep_pmax(L.h);
    if (last_ms) *last_ms = level < (int)s->level_ms.size() ? s->level_ms[level] : 0.0;
    return FRH_OK;
}

int frh_solver_set_pass_limit(frh_solver *s, int level, int pass_limit)
{
    FRH_REQ(s && level >= 0 && level < s->nl, "frh_solver_set_pass_limit: bad level");
    LevelDev &L = s->lv[level];
    FRH_REQ(pass_limit >= 1 && pass_limit <= sweep_pmax(L.h),
            "frh_solver_set_pass_limit: %d sweeps per pass outside [1, %d] for a %d-row level", pass_limit,
            sweep_pmax(L.h), L.h);
    L.passes = plan_passes(s->iterations, s->update_lag, pass_limit);
    if (s->gexec) {  // the captured level loop holds the old pass schedule
        (void)hipGraphExecDestroy(s->gexec);
        s->gexec = nullptr;
    }
    return FRH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- work list
// Uneven batches: several solvers (one per frame size and channel count,
// each holding its own workspace, weight levels and initial flow) run the
// level loop of frh_solver_run_pairs in lock-step.  Their pyramids are
// aligned at the finest level (level step k holds level k - (K - levels) of
// a solver with `levels` levels, K the deepest pyramid), and every stage of a
// level step is one launch over the frames of all solvers at that step: the
// level images, the flow initialisation or upsampling, the warp, the motion
// tensor, each pass's assembly and sweep, and the median.  The passes follow
// one plan per kernel form: levels of at most 512 rows in 512-lane
// work-groups (5 sweeps per pass), levels of 513 .. 1024 rows in 1024-lane
// work-groups (2 sweeps per pass) on a second stream.
struct frh_worklist {
    std::vector<frh_solver *> g;
    int K = 0;                       // level steps
    std::vector<Pass> plan512, plan1024;
    WlGroup *dev = nullptr;          // descriptors of every step (device)
    WlGroup *host = nullptr;         // their staging (pinned host memory)
    size_t cap = 0;                  // descriptors the buffers hold
    hipEvent_t staged = nullptr;     // the last descriptor upload has been read
    hipEvent_t done = nullptr;       // the last run's launches have finished
    bool pending = false;
    hipStream_t side = nullptr;      // 1024-lane passes
    hipEvent_t fork = nullptr, join = nullptr, first = nullptr;
};

static void free_worklist(frh_worklist *wl)
{
    if (!wl) return;
    if (wl->pending && wl->staged) (void)hipEventSynchronize(wl->staged);
    (void)hipFree(wl->dev);
    (void)hipHostFree(wl->host);
    if (wl->staged) (void)hipEventDestroy(wl->staged);
    if (wl->done) (void)hipEventDestroy(wl->done);
    if (wl->fork) (void)hipEventDestroy(wl->fork);
    if (wl->join) (void)hipEventDestroy(wl->join);
    if (wl->first) (void)hipEventDestroy(wl->first);
    if (wl->side) (void)hipStreamDestroy(wl->side);
    delete wl;
}

// One launch's descriptor array: its groups and item prefix sums.
struct WlLaunch {
    size_t off = 0;   // first descriptor in the step buffers
    WlStarts s{};
    int max_h = 0, max_w = 0, max_hs = 0, max_tiles = 0, max_median = 0;
    bool any_warp = false;
};

extern "C" {

int frh_worklist_create(frh_worklist **out, frh_solver *const *solvers, int n)
{
    FRH_REQ(out && solvers, "frh_worklist_create: null pointer");
    FRH_REQ(n >= 1 && n <= kWlMaxGroups, "frh_worklist_create: %d solvers outside [1, %d]", n, kWlMaxGroups);
    const frh_solver *s0 = solvers[0];
    FRH_REQ(s0, "frh_worklist_create: null solver");
    for (int k = 0; k < n; ++k) {
        const frh_solver *s = solvers[k];
        FRH_REQ(s, "frh_worklist_create: null solver %d", k);
        FRH_REQ(s->iterations == s0->iterations && s->update_lag == s0->update_lag,
                "frh_worklist_create: solvers differ in iterations or update_lag");
        FRH_REQ(s->const_assumption == s0->const_assumption, "frh_worklist_create: solvers differ in data term");
        FRH_REQ(s->a_smooth == 1.0, "frh_worklist_create: a_smooth != 1 is not supported in a work list");
        FRH_REQ(s->C <= 4, "frh_worklist_create: more than 4 channels");
        for (int i = 0; i < s->nl; ++i) {
            const LevelDev &L = s->lv[i];
            FRH_REQ(L.sk.hs <= 1024, "frh_worklist_create: a level of %d rows (at most 1024 in a work list)", L.h);
            // the sweep's buffer offsets are 32-bit
            FRH_REQ(L.sk.plane * sizeof(double) * kCoefPlanes < 0x80000000ull,
                    "frh_worklist_create: a %d x %d level is too large for 32-bit offsets", L.h, L.w);
            FRH_REQ(i == 0 || L.P_up == 4, "frh_worklist_create: flow upsampling with %d taps", L.P_up);
        }
        for (int j = 0; j < k; ++j) FRH_REQ(solvers[j] != s, "frh_worklist_create: solver %d listed twice", k);
    }
    frh_worklist *wl = new frh_worklist();
    for (int k = 0; k < n; ++k) {
        wl->g.push_back(solvers[k]);
        wl->K = std::max(wl->K, solvers[k]->nl);
    }
    wl->plan512 = plan_passes(s0->iterations, s0->update_lag, 5);
    wl->plan1024 = plan_passes(s0->iterations, s0->update_lag, 2);
    // per step: all active groups, those of the 512-lane form, those of the 1024-lane form
    wl->cap = (size_t)3 * wl->K * n;
    hipError_t e = hipMalloc((void **)&wl->dev, wl->cap * sizeof(WlGroup));
    if (e == hipSuccess) e = hipHostMalloc((void **)&wl->host, wl->cap * sizeof(WlGroup), hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&wl->staged, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&wl->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&wl->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&wl->join, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&wl->first, hipEventDisableTiming);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&wl->side, hipStreamNonBlocking);
    if (e != hipSuccess) {
        free_worklist(wl);
        return fail(FRH_EHIP, "frh_worklist_create: %s", hipGetErrorString(e));
    }
    *out = wl;
    return FRH_OK;
}

void frh_worklist_destroy(frh_worklist *wl) { free_worklist(wl); }

int frh_worklist_run(frh_worklist *wl, const float *const *fixed, const float *const *moving, const int *T,
                     void *const *w_out, int out_f64, void *stream)
{
    FRH_REQ(wl && fixed && moving && T && w_out, "frh_worklist_run: null pointer");
    const int n = (int)wl->g.size();
    hipStream_t st = as_stream(stream);
    for (int k = 0; k < n; ++k) {
        frh_solver *s = wl->g[k];
        FRH_REQ(fixed[k] && moving[k] && w_out[k], "frh_worklist_run: null pointer for solver %d", k);
        FRH_REQ(s->have_ref, "frh_worklist_run: frh_solver_set_reference has not been called for solver %d", k);
        FRH_REQ(T[k] >= 1 && T[k] <= s->max_batch, "frh_worklist_run: T=%d outside [1, %d] for solver %d", T[k],
                s->max_batch, k);
        FRH_REQ(!s->wlow || out_f64, "frh_worklist_run: min_level > 0 needs float64 output");
        if (!s->pyr1) {
            size_t T_ = (size_t)s->max_batch, C = s->C;
            int rc = dev_alloc(&s->pyr1, T_ * C * (size_t)s->H * s->Wsum, "reference pyramid rows");
            if (!rc) rc = dev_alloc(&s->f1l, T_ * C * (size_t)s->lv.back().h * s->lv.back().w, "reference levels");
            if (rc) return rc;
        }
    }
    // descriptors of every step, staged in pinned memory: the previous run's
    // upload must have been read before they are overwritten, and its
    // kernels must have finished before the device copy is replaced
    if (wl->pending) {
        FRH_HIP(hipEventSynchronize(wl->staged), "work list staging");
        FRH_HIP(hipStreamWaitEvent(st, wl->done, 0), "work list order");
    }
    const int K = wl->K;
    std::vector<WlLaunch> all(K), f512(K), f1024(K);
    size_t used = 0;
    auto add = [&](WlLaunch &L, const WlGroup &G) {
        if (L.s.n == 0) L.off = used;
        wl->host[used++] = G;
        L.s.start[L.s.n + 1] = L.s.start[L.s.n] + G.T;
        L.s.n += 1;
        L.max_h = std::max(L.max_h, G.h);
        L.max_w = std::max(L.max_w, G.w);
        L.max_hs = std::max(L.max_hs, G.hs);
        L.max_tiles = std::max(L.max_tiles, ((G.w + 31) / 32) * ((G.h + 15) / 16));
        L.max_median = std::max(L.max_median, (G.hs / 64) * (G.h + G.w - 1));
        L.any_warp = L.any_warp || G.li > 0;
    };
    // within a launch, the groups with the largest levels come first, so
    // that the work-groups of the longest frames are dispatched first
    std::vector<int> order(n);
    for (int pass = 0; pass < 3; ++pass) {
        for (int k = 0; k < K; ++k) {
            for (int gi = 0; gi < n; ++gi) order[gi] = gi;
            auto cells = [&](int gi) {
                const frh_solver *s = wl->g[gi];
                const int li = k - (K - s->nl);
                return li < 0 ? 0.0 : (double)s->lv[li].h * s->lv[li].w;
            };
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells(a) > cells(b); });
            for (int gi : order) {
                frh_solver *s = wl->g[gi];
                const int li = k - (K - s->nl);
                if (li < 0) continue;
                const LevelDev &L = s->lv[li];
                if (pass == 1 && L.sk.hs > 512) continue;
                if (pass == 2 && L.sk.hs <= 512) continue;
                WlGroup G{};
                G.T = T[gi];
                G.C = s->C;
                G.li = li;
                G.h = L.h; G.w = L.w; G.hs = L.sk.hs;
                G.plane = L.sk.plane;
                G.fs = s->npl * L.sk.plane;
                G.H = s->H; G.Wsum = s->Wsum; G.xoff = s->xoff[li];
                G.P_img = L.P_img; G.img_iy = L.img_iy; G.img_wy = L.img_wy;
                if (li > 0) {
                    G.ph = s->lv[li - 1].h; G.pw = s->lv[li - 1].w;
                    G.up_ix = L.up_ix; G.up_iy = L.up_iy; G.up_wx = L.up_wx; G.up_wy = L.up_wy;
                }
                G.h_row = L.h_row; G.h_col = L.h_col; G.axw = L.axw; G.ayw = L.ayw;
                G.pyr = s->pyr; G.pyr1 = s->pyr1;
                G.f2l = s->f2l; G.f1l = s->f1l; G.f2w = s->f2w; G.uvl = s->uvl; G.uvn = s->uvn;
                G.uvtop = s->uvtop;
                G.Wlev = s->Wlev + L.W_off;
                G.work = s->work;
                G.do_median = (L.h < L.w ? L.h : L.w) > 5;
                if (li + 1 < s->nl) G.out_kind = 0;
                else if (s->wlow) { G.out_kind = 3; G.out = s->wlow; }
                else { G.out_kind = out_f64 ? 2 : 1; G.out = w_out[gi]; }
                G.ad = make_adata(s->C, s->a_data.data(), s->am1.data(), -1.0);
                add(pass == 0 ? all[k] : (pass == 1 ? f512[k] : f1024[k]), G);
            }
        }
    }
    FRH_HIP(hipMemcpyAsync(wl->dev, wl->host, used * sizeof(WlGroup), hipMemcpyHostToDevice, st),
            "work list descriptors");
    FRH_HIP(hipEventRecord(wl->staged, st), "work list staging");
    wl->pending = true;
    // pyramid rows of every frame (per solver: its frame size)
    for (int gi = 0; gi < n; ++gi) {
        frh_solver *s = wl->g[gi];
        const int NC = T[gi] * s->C;
        FRH_HIP(launch_pyr_h(moving[gi], NC, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr, st),
                "moving pyramid rows");
        FRH_HIP(launch_pyr_h(fixed[gi], NC, s->H, s->W, s->Wsum, s->hoff, s->hix, s->hwx, s->pyr1, st),
                "reference pyramid rows");
    }
    const int kind = wl->g[0]->const_assumption;
    for (int k = 0; k < K; ++k) {
        const WlLaunch &A = all[k];
        const WlGroup *ga = wl->dev + A.off;
        FRH_HIP(launch_wl_pyr_v(ga, A.s, A.max_h, A.max_w, st), "work list: level images");
        FRH_HIP(launch_wl_flow(ga, A.s, A.max_h, A.max_w, st), "work list: level flow");
        if (A.any_warp) FRH_HIP(launch_wl_warp(ga, A.s, A.max_h, A.max_w, st), "work list: warp");
        FRH_HIP(launch_wl_tensor(kind, ga, A.s, A.max_tiles, kCensusEps, st), "work list: motion tensor");
        // The 1024-lane passes (the step's longest chain) go first, on the
        // second stream; the 512-lane passes start after its first
        // assembly, so that their work-groups fill the CUs that the
        // shorter tall frames leave idle during the 1024-lane sweeps.
        const bool tall = f1024[k].s.n > 0;
        if (tall) {
            FRH_HIP(hipEventRecord(wl->fork, st), "work list fork");
            FRH_HIP(hipStreamWaitEvent(wl->side, wl->fork, 0), "work list fork");
        }
        for (int form = 1; form >= 0; --form) {
            const WlLaunch &F = form == 0 ? f512[k] : f1024[k];
            if (F.s.n == 0) continue;
            const std::vector<Pass> &plan = form == 0 ? wl->plan512 : wl->plan1024;
            hipStream_t fs = form == 0 ? st : wl->side;
            const WlGroup *gf = wl->dev + F.off;
            for (size_t p = 0; p < plan.size(); ++p) {
                if (plan[p].assemble)
                    FRH_HIP(launch_wl_assemble(gf, F.s, F.max_h + F.max_w - 1, F.max_hs, plan[p].psi_one ? 1 : 0,
                                               p == 0 ? 1 : 0, fs), "work list: assemble");
                if (form == 1 && p == 0) {
                    FRH_HIP(hipEventRecord(wl->first, wl->side), "work list first assembly");
                    FRH_HIP(hipStreamWaitEvent(st, wl->first, 0), "work list first assembly");
                }
                FRH_HIP(launch_wl_sweep(plan[p].nsweeps, form, gf, F.s, p == 0 ? 1 : 0, fs), "work list: sweep");
            }
        }
        if (tall) {
            FRH_HIP(hipEventRecord(wl->join, wl->side), "work list join");
            FRH_HIP(hipStreamWaitEvent(st, wl->join, 0), "work list join");
        }
        FRH_HIP(launch_wl_median(ga, A.s, A.max_median, st), "work list: median");
    }
    // min_level > 0: cv2.resize of the coarse float64 flow to full resolution
    for (int gi = 0; gi < n; ++gi) {
        frh_solver *s = wl->g[gi];
        if (!s->wlow) continue;
        const LevelDev &L = s->lv.back();
        FRH_HIP(launch_cv_resize(s->wlow, T[gi], L.h, L.w, 2, s->wtmp, (double *)w_out[gi], s->H, s->W, st),
                "final cubic resize");
    }
    FRH_HIP(hipEventRecord(wl->done, st), "work list done");
    return FRH_OK;
}

}  // extern "C"
