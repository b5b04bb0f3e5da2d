// C ABI of libhfx: plan construction (validation, device tables, workspaces) and the
// launch sequences of the fused Doppler-moment path and the per-node kinds.
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "../../include/hfx.h"
#include "hfx_kernels.hh"

namespace {

thread_local std::string g_err;

}  // namespace

namespace hfx {
void set_last_error(const std::string &msg) { g_err = msg; }
}  // namespace hfx

namespace {

hfx_status fail(hfx_status code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HFX_TRY_HIP(expr)                                                              \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(HFX_E_HIP, std::string(#expr " failed: ") + hipGetErrorString(e_)); \
  } while (0)

inline hipStream_t S(hfx_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

// Device allocation that records itself for release.
struct DevArena {
  std::vector<void *> ptrs;
  size_t bytes = 0;
  template <typename T> hfx_status alloc(T **p, size_t count) {
    void *v = nullptr;
    size_t b = count * sizeof(T);
    if (b == 0) b = 16;
    hipError_t e = hipMalloc(&v, b);
    if (e != hipSuccess) return fail(HFX_E_NOMEM, "hipMalloc of " + std::to_string(b) + " bytes failed");
    ptrs.push_back(v);
    bytes += b;
    *p = static_cast<T *>(v);
    return HFX_OK;
  }
  ~DevArena() {
    for (void *p : ptrs) (void)hipFree(p);
  }
};

#define HFX_TRY(expr)              \
  do {                             \
    hfx_status s_ = (expr);        \
    if (s_ != HFX_OK) return s_;   \
  } while (0)

// M1/M2 weights exactly as the reference builds them: Arange in double, cast to F32
// (arange.cu:61-68), negative band at f - fs (graph_builder.cc:578-601), squares in F32
// (Multiply of two F32 tensors, graph_builder.cc:626).
void build_weights(int kb, int ke, int T, double fs_w, bool both, std::vector<float> &f,
                   std::vector<float> &f2) {
  const double df = fs_w / (double)T;
  const double f0 = kb * df;
  const int nb = ke - kb;
  f.clear();
  for (int i = 0; i < nb; ++i) f.push_back((float)(f0 + df * (double)i));
  if (both)
    for (int i = 0; i < nb; ++i) f.push_back((float)((f0 - fs_w) + df * (double)i));
  f2.resize(f.size());
  for (size_t i = 0; i < f.size(); ++i) f2[i] = f[i] * f[i];
}

}  // namespace

// A spectral filter: forward row pass -> column pass with the table between its forward
// and inverse transforms -> inverse row pass (Filter2D, and Flatfield's Fourier form).
struct SpectralFilter {
  int64_t batch = 0;
  int H = 0, W = 0, blk_c = 1;
  DevArena arena;
  float2 *tw_w = nullptr, *tw_h = nullptr, *tf = nullptr;
  float2 *ws = nullptr;  // [batch][H][W] blocked spectrum: owned, or borrowed from a plan
};

struct hfx_plan {
  hfx_plan_desc d{};
  int H = 0, W = 0, OH = 0, OW = 0;
  long long npix = 0, onpix = 0;
  int nb = 0, nbc = 0, n_planes = 0, n_out = 3, blk_c = 1, wm_rows = 0;
  bool fused_row = false;  // time DFT fused with the forward row FFT
  DevArena arena;
  float2 *planes_nat = nullptr, *planes_blk = nullptr;
  float2 *tw_w = nullptr, *tw_h = nullptr;
  float *wm32 = nullptr;
  __half *wm16hi = nullptr, *wm16lo = nullptr;
  __hip_bfloat16 *wmbf = nullptr;
  float2 *lx = nullptr, *ly = nullptr;
  float2 *tf = nullptr;
  float *wf = nullptr, *wf2 = nullptr;
  int nwin = 1;              // windows per front launch in hfx_moments_batch
  int lanes = 1;             // streams hfx_moments_batch alternates over (1 or 2)
  hipStream_t aux = nullptr; // second lane's stream (default priority, non-blocking)
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  bool filter2d = false;     // Fresnel: Filter2D on the band planes before the lens
  SpectralFilter f2d;        // its tables; workspace = planes_blk
  hfx_flatfield *ff = nullptr;  // Flatfield on the moment images
  float *mom_tmp = nullptr;  // moment images before the flatfield
  ~hfx_plan();
};

struct hfx_fresnel {
  long long batch = 0;
  int H = 0, W = 0, in_real = 0, output_magnitude = 0, skip_phase_shift = 0;
  double lambda = 0, dx = 0, dy = 0, z = 0;
  int blk_c = 1;
  DevArena arena;
  float2 *tw_w = nullptr, *tw_h = nullptr, *lx = nullptr, *ly = nullptr, *ws = nullptr;
};

struct hfx_asm {
  hfx_asm_desc d{};
  int OH = 0, OW = 0, blk_c = 1;
  DevArena arena;
  float2 *tw_w = nullptr, *tw_h = nullptr, *tf = nullptr, *ws = nullptr;
};

struct hfx_filter2d {
  SpectralFilter sf;
  int in_real = 0;
};

struct hfx_flatfield {
  int64_t batch = 0;
  int H = 0, W = 0;
  bool use_fft = false;
  int rx = 0, ry = 0;
  DevArena arena;
  float *kx = nullptr, *ky = nullptr, *tmp = nullptr;
  SpectralFilter sf;
};

hfx_plan::~hfx_plan() {
  if (ff) (void)hfx_flatfield_destroy(ff);
  if (aux) (void)hipStreamSynchronize(aux);
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  if (ev_join) (void)hipEventDestroy(ev_join);
  if (aux) (void)hipStreamDestroy(aux);
}

static hfx_status spectral_setup(SpectralFilter &f, int64_t batch, int H, int W, hipStream_t st,
                                 float2 *borrowed_ws = nullptr);
static hfx_status spectral_run(const SpectralFilter &f, const void *in, bool in_real, void *out, bool out_real,
                               hipStream_t st);
static hfx_status window_back_moments(hfx_plan *p, float2 *planes, float *out, hipStream_t st);

extern "C" {

const char *hfx_last_error(void) { return g_err.c_str(); }
int32_t hfx_abi_version(void) { return 1; }

void hfx_plan_desc_init(hfx_plan_desc *d) {
  std::memset(d, 0, sizeof(*d));
  d->height = 512;
  d->width = 512;
  d->window = 32;
  d->stride = 32;
  d->fs_hz = 37e3;
  d->lambda_m = 852e-9;
  d->dx_m = 12e-6;
  d->dy_m = 12e-6;
  d->z_m = 0.3;
  d->bin_begin = 6;
  d->bin_end = 13;
  d->prop = HFX_PROP_FRESNEL;
  d->precision = HFX_PREC_FP32;
  d->flags = 0;
  d->weight_fs_hz = 37e3;
}

static hfx_status gen_prop_tables(hfx_plan *p, hipStream_t st) {
  const hfx_plan_desc &d = p->d;
  if (d.prop == HFX_PROP_FRESNEL) {
    hfx::launch_fresnel_chirp(p->lx, p->ly, p->H, p->W, d.lambda_m, d.z_m, d.dx_m, d.dy_m, st);
  } else {
    hfx::launch_asm_transfer(p->tf, p->OH, p->OW, d.lambda_m, d.z_m, d.dx_m, d.dy_m,
                             d.filter_enabled, d.r_inner, d.r_outer, d.s_inner, d.s_outer, p->blk_c,
                             st);
  }
  HFX_TRY_HIP(hipGetLastError());
  return HFX_OK;
}

hfx_status hfx_plan_create(const hfx_plan_desc *dp, hfx_stream_t stream, hfx_plan **out) {
  if (!dp || !out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  const hfx_plan_desc &d = *dp;
  if (d.height <= 0 || d.width <= 0) return fail(HFX_E_INVALID, "frame size must be positive");
  if (d.window <= 0 || d.stride <= 0) return fail(HFX_E_INVALID, "window and stride must be positive");
  if (d.window % 4 != 0) return fail(HFX_E_UNSUPPORTED, "window must be a multiple of 4");
  if (d.precision != HFX_PREC_FP32 && d.window % 32 != 0)
    return fail(HFX_E_UNSUPPORTED, "16-bit MFMA time DFT needs a window multiple of 32");
  if (d.precision < 0 || d.precision > 3) return fail(HFX_E_INVALID, "unknown precision");
  if (d.width % 4 != 0) return fail(HFX_E_UNSUPPORTED, "frame width must be a multiple of 4");
  if ((long long)d.height * d.width > INT32_MAX) return fail(HFX_E_UNSUPPORTED, "frame holds more than 2^31 - 1 pixels");
  if (!(d.lambda_m > 0)) return fail(HFX_E_INVALID, "wavelength must be positive");
  if (!(d.dx_m > 0) || !(d.dy_m > 0)) return fail(HFX_E_INVALID, "pixel pitch must be positive");
  if (d.dx_m != d.dy_m) return fail(HFX_E_INVALID, "dx must equal dy");
  const bool both = (d.flags & HFX_FLAG_BOTH_SIDEBANDS) != 0;
  const int half = d.window / 2 + 1;
  if (d.bin_begin < 0 || d.bin_end <= d.bin_begin)
    return fail(HFX_E_INVALID, "band must satisfy 0 <= bin_begin < bin_end");
  if (!both && d.bin_end > half)
    return fail(HFX_E_INVALID, "bin_end must be <= window/2 + 1 for the one-sided DFT");
  if (both && d.bin_end > d.window) return fail(HFX_E_INVALID, "bin_end must be <= window");
  if (d.prop != HFX_PROP_FRESNEL && d.prop != HFX_PROP_ANGULAR_SPECTRUM)
    return fail(HFX_E_INVALID, "unknown propagation");
  if (d.prop == HFX_PROP_FRESNEL && !(d.z_m != 0)) return fail(HFX_E_INVALID, "Fresnel z must be non-zero");
  if (d.prop == HFX_PROP_FRESNEL && (d.pad_height || d.pad_width))
    return fail(HFX_E_INVALID, "padding applies to the angular spectrum only");
  if (d.filter_enabled && !(d.r_inner >= 0 && d.r_outer >= 0 && d.s_inner >= 0 && d.s_outer >= 0))
    return fail(HFX_E_INVALID, "filter radii and tapers must be non-negative");
  if (d.filter_enabled && d.prop == HFX_PROP_FRESNEL &&
      (d.r_inner != std::floor(d.r_inner) || d.r_outer != std::floor(d.r_outer) ||
       d.s_inner != std::floor(d.s_inner) || d.s_outer != std::floor(d.s_outer) || d.r_outer > 1e9 ||
       d.r_inner > 1e9 || d.s_inner > 1e9 || d.s_outer > 1e9))
    return fail(HFX_E_INVALID, "Filter2D radii and tapers must be integers");
  if (d.flatfield_enabled && !(d.flatfield_sigma_y > 0.0f && d.flatfield_sigma_x > 0.0f))
    return fail(HFX_E_INVALID, "flatfield sigma must be positive");
  const int nb = d.bin_end - d.bin_begin;
  if (d.split_bin != 0 && (d.split_bin <= d.bin_begin || d.split_bin >= d.bin_end || both))
    return fail(HFX_E_INVALID, "split_bin must lie strictly inside the one-sided band");

  auto *p = new (std::nothrow) hfx_plan();
  if (!p) return fail(HFX_E_NOMEM, "host allocation failed");
  std::unique_ptr<hfx_plan> guard(p);
  p->d = d;
  p->H = d.height;
  p->W = d.width;
  p->OH = d.pad_height > 0 ? d.pad_height : d.height;
  p->OW = d.pad_width > 0 ? d.pad_width : d.width;
  if (p->OH < p->H || p->OW < p->W || (p->OH - p->H) % 2 || (p->OW - p->W) % 2)
    return fail(HFX_E_INVALID, "padding must be >= the frame size with even margins");
  p->npix = (long long)p->H * p->W;
  p->onpix = (long long)p->OH * p->OW;
  p->nb = nb;
  p->nbc = both ? nb + 1 : nb;
  if (p->nbc > 64) return fail(HFX_E_UNSUPPORTED, "at most 64 band bins per plan");
  p->n_planes = both ? 2 * nb : nb;
  p->n_out = d.split_bin > 0 ? 5 : 3;
  p->blk_c = hfx::pick_blk_c(p->OW, p->OH);
  if (!hfx::row_pass_supported(p->OW) || !hfx::col_pass_supported(p->OH, p->blk_c))
    return fail(HFX_E_UNSUPPORTED, "propagation size must be one of the compiled 2^a*3^b lengths "
                                   "(2..2048)");
  p->wm_rows = ((2 * p->nbc + 15) / 16) * 16;

  hipStream_t st = S(stream);
  DevArena &A = p->arena;
  p->filter2d = d.prop == HFX_PROP_FRESNEL && d.filter_enabled;
  p->fused_row = !p->filter2d && p->OH == p->H && p->OW == p->W && hfx::dft_row_supported(p->W, p->nbc, d.window);
  if (d.batch_windows > 1 && p->fused_row &&
      hfx::dft_front_pipelined(p->W, p->nbc, d.window, p->n_planes, d.precision))
    p->nwin = std::min<int>(d.batch_windows, hfx::HFX_MAX_WIN);
  // batch lanes: a second stream (and plane workspaces) lets one window's kernels overlap the
  // neighbouring window's (measured +15 % at 1024^2, T = 64, in FP32 and the 16-bit modes);
  // the Flatfield's scratch images and Filter2D's tables are shared, so those keep one lane.
  if (d.batch_lanes < 0 || d.batch_lanes > 2) return fail(HFX_E_INVALID, "batch_lanes must be 0, 1 or 2");
  p->lanes = d.batch_lanes != 1 && !p->filter2d && !d.flatfield_enabled ? 2 : 1;
  // natural-layout planes, one set per lane: only the stand-alone time DFT needs them
  HFX_TRY(A.alloc(&p->planes_nat, p->fused_row ? 1 : (size_t)p->lanes * p->n_planes * p->npix));
  if (p->lanes == 2) {
    HFX_TRY_HIP(hipStreamCreateWithFlags(&p->aux, hipStreamNonBlocking));
    HFX_TRY_HIP(hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming));
    HFX_TRY_HIP(hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming));
  }
  // one blocked plane workspace per window of a front launch, per lane
  HFX_TRY(A.alloc(&p->planes_blk, (size_t)p->lanes * p->nwin * p->n_planes * p->onpix));
  HFX_TRY(A.alloc(&p->tw_w, p->OW));
  HFX_TRY(A.alloc(&p->tw_h, p->OH));
  HFX_TRY(A.alloc(&p->wm32, (size_t)p->wm_rows * d.window));
  HFX_TRY(A.alloc(&p->wm16hi, (size_t)p->wm_rows * d.window));
  HFX_TRY(A.alloc(&p->wm16lo, (size_t)p->wm_rows * d.window));
  HFX_TRY(A.alloc(&p->wmbf, (size_t)p->wm_rows * d.window));
  if (d.prop == HFX_PROP_FRESNEL) {
    HFX_TRY(A.alloc(&p->lx, p->W));
    HFX_TRY(A.alloc(&p->ly, p->H));
  } else {
    HFX_TRY(A.alloc(&p->tf, (size_t)p->onpix));
  }
  HFX_TRY(A.alloc(&p->wf, p->n_planes));
  HFX_TRY(A.alloc(&p->wf2, p->n_planes));

  hfx::launch_fft_twiddles(p->tw_w, p->OW, st);
  hfx::launch_fft_twiddles(p->tw_h, p->OH, st);
  hfx::launch_dft_matrix(p->wm32, p->wm16hi, p->wm16lo, p->wmbf, p->wm_rows, d.window,
                         d.bin_begin, p->nbc, st);
  HFX_TRY(gen_prop_tables(p, st));
  if (p->filter2d) {
    HFX_TRY(spectral_setup(p->f2d, p->n_planes, p->H, p->W, st, p->planes_blk));
    hfx::launch_filter2d_table(p->f2d.tf, p->H, p->W, (int)d.r_inner, (int)d.r_outer, (int)d.s_inner,
                               (int)d.s_outer, 1, p->f2d.blk_c, st);
  }
  if (d.flatfield_enabled) {
    HFX_TRY(A.alloc(&p->mom_tmp, (size_t)p->n_out * p->onpix));
    HFX_TRY(hfx_flatfield_create(p->n_out, p->OH, p->OW, d.flatfield_sigma_y, d.flatfield_sigma_x, stream, &p->ff,
                                 nullptr));
  }
  std::vector<float> f, f2;
  double fs_w = d.weight_fs_hz > 0 ? d.weight_fs_hz : d.fs_hz;
  build_weights(d.bin_begin, d.bin_end, d.window, fs_w, both, f, f2);
  HFX_TRY_HIP(hipMemcpyAsync(p->wf, f.data(), f.size() * 4, hipMemcpyHostToDevice, st));
  HFX_TRY_HIP(hipMemcpyAsync(p->wf2, f2.data(), f2.size() * 4, hipMemcpyHostToDevice, st));
  HFX_TRY_HIP(hipStreamSynchronize(st));
  *out = guard.release();
  return HFX_OK;
}

hfx_status hfx_plan_destroy(hfx_plan *p) {
  delete p;
  return HFX_OK;
}

hfx_status hfx_plan_get_info(const hfx_plan *p, hfx_plan_info *info) {
  if (!p || !info) return fail(HFX_E_INVALID, "null argument");
  info->out_height = p->OH;
  info->out_width = p->OW;
  info->n_outputs = p->n_out;
  info->n_planes = p->n_planes;
  info->workspace_bytes = (int64_t)(p->arena.bytes + p->f2d.arena.bytes);
  info->batch_lanes = p->lanes;
  info->groupable = p->fused_row && !p->filter2d && !p->ff &&
                    hfx::dft_front_pipelined(p->W, p->nbc, p->d.window, p->n_planes, p->d.precision) &&
                    hfx::wave_pass_supported(p->OH) && hfx::wave_pass_supported(p->OW);
  return HFX_OK;
}

hfx_status hfx_plan_update_z(hfx_plan *p, double z_m, hfx_stream_t stream) {
  if (!p) return fail(HFX_E_INVALID, "null plan");
  if (p->d.prop == HFX_PROP_FRESNEL && !(z_m != 0)) return fail(HFX_E_INVALID, "Fresnel z must be non-zero");
  p->d.z_m = z_m;
  HFX_TRY(gen_prop_tables(p, S(stream)));
  return HFX_OK;
}

static hfx_status check_ring(const hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                             int64_t first) {
  if (!p || !ring) return fail(HFX_E_INVALID, "null argument");
  if (ring_frames < p->d.window) return fail(HFX_E_INVALID, "ring holds fewer frames than the window");
  if (ring_frames > INT32_MAX) return fail(HFX_E_UNSUPPORTED, "ring holds more than 2^31 - 1 frames");
  if (first < 0 || first >= ring_frames) return fail(HFX_E_INVALID, "first_frame outside the ring");
  if ((reinterpret_cast<uintptr_t>(ring) & 7) != 0) return fail(HFX_E_INVALID, "ring must be 8-byte aligned");
  return HFX_OK;
}

static hfx::DftLaunch dft_launch(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                                 int64_t first, float2 *planes) {
  hfx::DftLaunch L{};
  L.ring = ring;
  L.ring_frames = ring_frames;
  L.first = first;
  L.T = p->d.window;
  L.npix = p->npix;
  L.W = p->W;
  L.kb = p->d.bin_begin;
  L.nb = p->nb;
  L.nbc = p->nbc;
  L.both = (p->d.flags & HFX_FLAG_BOTH_SIDEBANDS) ? 1 : 0;
  L.planes = planes;
  // the Fresnel lens multiplies the DFT output, unless Filter2D comes between them
  L.lx = p->d.prop == HFX_PROP_FRESNEL && !p->filter2d ? p->lx : nullptr;
  L.ly = p->d.prop == HFX_PROP_FRESNEL && !p->filter2d ? p->ly : nullptr;
  L.precision = p->d.precision;
  L.wm_f32 = p->wm32;
  L.wm_f16hi = reinterpret_cast<const uint16_t *>(p->wm16hi);
  L.wm_f16lo = reinterpret_cast<const uint16_t *>(p->wm16lo);
  L.wm_bf16 = reinterpret_cast<const uint16_t *>(p->wmbf);
  return L;
}

static hfx_status run_dft(hfx_plan *p, const uint16_t *ring, int64_t ring_frames, int64_t first,
                          float2 *planes, hipStream_t st) {
  int rc = hfx::launch_time_dft(dft_launch(p, ring, ring_frames, first, planes), st);
  if (rc != 0) return fail(HFX_E_HIP, "time DFT launch failed (" + std::to_string(rc) + ")");
  return HFX_OK;
}

hfx_status hfx_plan_time_dft(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                             int64_t first_frame, void *planes_out, hfx_stream_t stream) {
  HFX_TRY(check_ring(p, ring, ring_frames, first_frame));
  if (!planes_out || (reinterpret_cast<uintptr_t>(planes_out) & 15) != 0)
    return fail(HFX_E_INVALID, "planes_out must be a 16-byte aligned device buffer");
  return run_dft(p, ring, ring_frames, first_frame, reinterpret_cast<float2 *>(planes_out),
                 S(stream));
}

static hfx::PassArgs base_pass(int H, int W, int planes, int blk_c) {
  hfx::PassArgs a;
  std::memset(&a, 0, sizeof(a));
  a.H = H;
  a.W = W;
  a.planes = planes;
  a.blk_c = blk_c;
  a.in_h = H;
  a.in_w = W;
  return a;
}

// Front of the schedule: the window's band planes, row-transformed, in the plan's blocked
// workspace (one fused kernel when the frame width allows, else time DFT + row pass).
// planes: the blocked workspace (default the plan's first); lane: the natural-layout planes'
// set of the stand-alone time DFT
static hfx_status window_front(hfx_plan *p, const uint16_t *ring, int64_t ring_frames, int64_t first,
                               hipStream_t st, float2 *planes = nullptr, int lane = 0) {
  if (!planes) planes = p->planes_blk;
  if (p->fused_row) {
    hfx::DftLaunch L = dft_launch(p, ring, ring_frames, first, planes);
    if (hfx::launch_time_dft_row(L, p->OH, p->blk_c, p->n_planes, p->tw_w, st) != 0)
      return fail(HFX_E_HIP, "fused time-DFT/row-FFT launch failed");
    return HFX_OK;
  }
  float2 *nat = p->planes_nat + (long long)lane * p->n_planes * p->npix;
  HFX_TRY(run_dft(p, ring, ring_frames, first, nat, st));
  // Filter2D in place on the natural planes, through the blocked workspace (one lane)
  if (p->filter2d) HFX_TRY(spectral_run(p->f2d, nat, false, nat, false, st));
  // row pass: natural planes (lens already applied for Fresnel, or here after Filter2D)
  // -> blocked spectrum rows
  hfx::PassArgs r = base_pass(p->OH, p->OW, p->n_planes, p->blk_c);
  if (p->filter2d) {
    r.pre_lx = p->lx;
    r.pre_ly = p->ly;
  }
  r.in = nat;
  r.in_mode = hfx::IN_NAT_C;
  r.in_h = p->H;
  r.in_w = p->W;
  r.off_y = (p->OH - p->H) / 2;
  r.off_x = (p->OW - p->W) / 2;
  r.out = planes;
  r.out_mode = hfx::OUT_BLK;
  r.tw = p->tw_w;
  if (hfx::launch_row_pass(r, st) != 0) return fail(HFX_E_HIP, "row pass launch failed");
  return HFX_OK;
}

// Back of the schedule: column pass (Fresnel: with |.| and the moment reduction; ASM: with
// the transfer function, then the inverse row pass with the reduction) into `out`.
static hfx_status window_back(hfx_plan *p, float *out, hipStream_t st, float2 *planes = nullptr) {
  HFX_TRY(window_back_moments(p, planes ? planes : p->planes_blk, p->ff ? p->mom_tmp : out, st));
  if (p->ff) return hfx_flatfield_execute(p->ff, p->mom_tmp, out, reinterpret_cast<hfx_stream_t>(st));
  return HFX_OK;
}

static hfx_status window_back_moments(hfx_plan *p, float2 *planes, float *out, hipStream_t st) {
  const long long on = p->onpix;
  hfx::MomentArgs mom{};
  for (int i = 0; i < p->n_out; ++i) mom.out[i] = out + i * on;
  mom.f = p->wf;
  mom.f2 = p->wf2;
  mom.split = p->d.split_bin > 0 ? p->d.split_bin - p->d.bin_begin : 0;
  mom.power2 = (p->d.flags & HFX_FLAG_POWER2) ? 1 : 0;
  mom.shift = (p->d.flags & HFX_FLAG_FFTSHIFT) ? 1 : 0;

  hfx::PassArgs c = base_pass(p->OH, p->OW, p->n_planes, p->blk_c);
  c.in = planes;
  c.in_mode = hfx::IN_BLK;
  c.tw = p->tw_h;
  if (p->d.prop == HFX_PROP_FRESNEL) {
    c.out_mode = hfx::OUT_MOMENTS;
    c.mom = mom;
    if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "column pass launch failed");
  } else {
    c.mid_tf = p->tf;
    c.out = planes;
    c.out_mode = hfx::OUT_BLK;
    if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "column pass launch failed");
    hfx::PassArgs r2 = base_pass(p->OH, p->OW, p->n_planes, p->blk_c);
    r2.in = planes;
    r2.in_mode = hfx::IN_BLK;
    r2.inv = 1;
    r2.out_mode = hfx::OUT_MOMENTS;
    r2.mom = mom;
    r2.tw = p->tw_w;
    if (hfx::launch_row_pass(r2, st) != 0) return fail(HFX_E_HIP, "inverse row pass launch failed");
  }
  return HFX_OK;
}

static hfx_status window_moments(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                                 int64_t first, float *out, hipStream_t st) {
  HFX_TRY(window_front(p, ring, ring_frames, first, st));
  return window_back(p, out, st);
}

// A group of plans whose windows can share one front launch and one launch per back pass:
// the same frame and output geometry, window, band, weights and kernels; the propagation
// tables (z, lambda, pitch), rings and outputs may differ.
static bool group_compatible(const hfx_plan *a, const hfx_plan *b) {
  const hfx_plan_desc &x = a->d, &y = b->d;
  return a->H == b->H && a->W == b->W && a->OH == b->OH && a->OW == b->OW && x.window == y.window &&
         x.prop == y.prop && x.precision == y.precision && x.flags == y.flags && x.bin_begin == y.bin_begin &&
         x.bin_end == y.bin_end && x.split_bin == y.split_bin && x.weight_fs_hz == y.weight_fs_hz &&
         x.pad_height == y.pad_height && x.pad_width == y.pad_width && a->n_planes == b->n_planes &&
         a->n_out == b->n_out && a->blk_c == b->blk_c && a->fused_row == b->fused_row && a->filter2d == b->filter2d &&
         (a->ff != nullptr) == (b->ff != nullptr);
}

hfx_status hfx_moments_group(hfx_plan *const *plans, const uint16_t *const *rings, const int64_t *ring_frames,
                             const int64_t *first_frames, float *const *outs, int32_t n, hfx_stream_t stream) {
  if (!plans || !rings || !ring_frames || !first_frames || !outs) return fail(HFX_E_INVALID, "null argument");
  static_assert(HFX_GROUP_MAX <= hfx::HFX_MAX_WIN, "grouped launches carry at most HFX_MAX_WIN windows");
  if (n < 1 || n > HFX_GROUP_MAX)
    return fail(HFX_E_INVALID, "a group holds 1 to " + std::to_string(HFX_GROUP_MAX) + " windows");
  hfx_plan *p0 = plans[0];
  for (int w = 0; w < n; ++w) {
    if (!plans[w] || !outs[w]) return fail(HFX_E_INVALID, "null plan or output in the group");
    HFX_TRY(check_ring(plans[w], rings[w], ring_frames[w], first_frames[w]));
    for (int v = 0; v < w; ++v)
      if (plans[v] == plans[w]) return fail(HFX_E_INVALID, "a plan appears twice in the group (one workspace each)");
    if (!group_compatible(p0, plans[w]))
      return fail(HFX_E_INVALID, "group plans differ in geometry, window, band, weights or options");
  }
  hipStream_t st = S(stream);
  if (n == 1) return window_moments(p0, rings[0], ring_frames[0], first_frames[0], outs[0], st);
  if (!p0->fused_row || p0->filter2d || p0->ff ||
      !hfx::dft_front_pipelined(p0->W, p0->nbc, p0->d.window, p0->n_planes, p0->d.precision) ||
      !hfx::wave_pass_supported(p0->OH) || !hfx::wave_pass_supported(p0->OW))
    return fail(HFX_E_UNSUPPORTED, "grouped windows need the pipelined FP32 front (width 512, 768 or 1024), "
                                   "wave-FFT sizes and no Filter2D or Flatfield");
  // front: one launch over the windows of all plans (each into its own plan's workspace)
  hfx::DftLaunch L = dft_launch(p0, rings[0], ring_frames[0], first_frames[0], p0->planes_blk);
  L.nwin = n;
  const bool fresnel = p0->d.prop == HFX_PROP_FRESNEL;
  for (int w = 0; w < n; ++w) {
    L.firsts[w] = first_frames[w];
    L.rings[w] = rings[w];
    L.ring_frames_w[w] = ring_frames[w];
    L.planes_w[w] = plans[w]->planes_blk;
    L.lx_w[w] = fresnel ? plans[w]->lx : nullptr;
    L.ly_w[w] = fresnel ? plans[w]->ly : nullptr;
  }
  if (hfx::launch_time_dft_row(L, p0->OH, p0->blk_c, p0->n_planes, p0->tw_w, st) != 0)
    return fail(HFX_E_HIP, "grouped front launch failed");
  // back: one launch per pass, blockIdx.z picking the window
  hfx::MomentArgs mom{};
  mom.f = p0->wf;
  mom.f2 = p0->wf2;
  mom.split = p0->d.split_bin > 0 ? p0->d.split_bin - p0->d.bin_begin : 0;
  mom.power2 = (p0->d.flags & HFX_FLAG_POWER2) ? 1 : 0;
  mom.shift = (p0->d.flags & HFX_FLAG_FFTSHIFT) ? 1 : 0;
  auto set_windows = [&](hfx::PassArgs &a, bool mid) {
    a.nwin = n;
    for (int w = 0; w < n; ++w) {
      a.win[w].in = plans[w]->planes_blk;
      a.win[w].out = mid ? plans[w]->planes_blk : nullptr;
      a.win[w].mid_tf = mid ? plans[w]->tf : nullptr;
      for (int i = 0; i < 5; ++i) a.win[w].mom_out[i] = i < p0->n_out ? outs[w] + i * p0->onpix : nullptr;
    }
    // the single-window fields stay consistent with window 0 (kinds are chosen from them)
    a.in = a.win[0].in;
    a.out = a.win[0].out;
    a.mid_tf = a.win[0].mid_tf;
    for (int i = 0; i < 5; ++i) a.mom.out[i] = a.win[0].mom_out[i];
  };
  hfx::PassArgs c = base_pass(p0->OH, p0->OW, p0->n_planes, p0->blk_c);
  c.in_mode = hfx::IN_BLK;
  c.tw = p0->tw_h;
  if (fresnel) {
    c.out_mode = hfx::OUT_MOMENTS;
    c.mom = mom;
    set_windows(c, false);
    if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "grouped column pass launch failed");
    return HFX_OK;
  }
  c.out_mode = hfx::OUT_BLK;
  set_windows(c, true);
  if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "grouped column pass launch failed");
  hfx::PassArgs r2 = base_pass(p0->OH, p0->OW, p0->n_planes, p0->blk_c);
  r2.in_mode = hfx::IN_BLK;
  r2.inv = 1;
  r2.out_mode = hfx::OUT_MOMENTS;
  r2.mom = mom;
  r2.tw = p0->tw_w;
  set_windows(r2, false);
  if (hfx::launch_row_pass(r2, st) != 0) return fail(HFX_E_HIP, "grouped inverse row pass launch failed");
  return HFX_OK;
}

hfx_status hfx_plan_run_stage(hfx_plan *p, int32_t stage, const uint16_t *ring, int64_t ring_frames,
                              int64_t first_frame, float *out, hfx_stream_t stream) {
  if (!p) return fail(HFX_E_INVALID, "null plan");
  if (stage == 0) {
    HFX_TRY(check_ring(p, ring, ring_frames, first_frame));
    return window_front(p, ring, ring_frames, first_frame, S(stream));
  }
  if (stage == 1) {
    if (!out) return fail(HFX_E_INVALID, "null output");
    return window_back(p, out, S(stream));
  }
  return fail(HFX_E_INVALID, "stage must be 0 (front) or 1 (back)");
}

hfx_status hfx_moments(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                       int64_t first_frame, float *out, hfx_stream_t stream) {
  HFX_TRY(check_ring(p, ring, ring_frames, first_frame));
  if (!out) return fail(HFX_E_INVALID, "null output");
  return window_moments(p, ring, ring_frames, first_frame, out, S(stream));
}

hfx_status hfx_moments_batch(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                             int64_t first_frame, int32_t n_windows, int64_t out_stride,
                             float *out, hfx_stream_t stream) {
  HFX_TRY(check_ring(p, ring, ring_frames, first_frame));
  if (!out || n_windows < 0) return fail(HFX_E_INVALID, "invalid output");
  if (out_stride < (int64_t)p->n_out * p->onpix) return fail(HFX_E_INVALID, "out_stride too small");
  hipStream_t st = S(stream);
  auto frame_of = [&](int w) { return (first_frame + (int64_t)w * p->d.stride) % ring_frames; };
  const long long pstride = (long long)p->n_planes * p->onpix;  // one window's planes
  // groups of nwin windows (one front launch each) alternate over the lanes; lane 1 runs on
  // the plan's own stream, forked from and joined back into the caller's stream
  const bool two = p->lanes == 2 && n_windows > p->nwin;
  if (two) {
    HFX_TRY_HIP(hipEventRecord(p->ev_fork, st));
    HFX_TRY_HIP(hipStreamWaitEvent(p->aux, p->ev_fork, 0));
  }
  hfx_status rc = HFX_OK;
  int g = 0;
  for (int w = 0; w < n_windows && rc == HFX_OK; w += p->nwin, ++g) {
    const int lane = two ? (g & 1) : 0;
    hipStream_t ls = lane ? p->aux : st;
    float2 *ws = p->planes_blk + (long long)lane * p->nwin * pstride;
    if (p->nwin > 1) {
      // nwin windows per front launch (the pipelined kernel runs on across their rows), then
      // the back of each window from its own plane workspace
      const int k = std::min(p->nwin, n_windows - w);
      hfx::DftLaunch L = dft_launch(p, ring, ring_frames, frame_of(w), ws);
      L.nwin = k;
      for (int j = 0; j < k; ++j) L.firsts[j] = frame_of(w + j);
      L.planes_stride = pstride;
      if (hfx::launch_time_dft_row(L, p->OH, p->blk_c, p->n_planes, p->tw_w, ls) != 0) {
        rc = fail(HFX_E_HIP, "multi-window front launch failed");
        break;
      }
      for (int j = 0; j < k && rc == HFX_OK; ++j) rc = window_back(p, out + (w + j) * out_stride, ls, ws + j * pstride);
    } else {
      rc = window_front(p, ring, ring_frames, frame_of(w), ls, ws, lane);
      if (rc == HFX_OK) rc = window_back(p, out + w * out_stride, ls, ws);
    }
  }
  if (two) {  // join even after a failed launch, so the caller's stream orders after lane 1
    const hipError_t e1 = hipEventRecord(p->ev_join, p->aux);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(st, p->ev_join, 0) : e1;
    if (rc == HFX_OK && e2 != hipSuccess) rc = fail(HFX_E_HIP, hipGetErrorString(e2));
  }
  return rc;
}

// ------------------------------------------------------------------------ per-node kinds

hfx_status hfx_dft_axis(const void *in, int32_t in_dtype, int64_t outer, int32_t n, int64_t inner,
                        int32_t k_begin, int32_t k_end, float scale, void *out, hfx_stream_t stream) {
  if (!in || !out) return fail(HFX_E_INVALID, "null argument");
  if (in_dtype < 1 || in_dtype > 3) return fail(HFX_E_INVALID, "unsupported input dtype");
  if (n <= 0 || outer < 0 || inner < 0) return fail(HFX_E_INVALID, "invalid shape");
  if (k_begin < 0 || k_end < k_begin || k_end > n) return fail(HFX_E_INVALID, "invalid bin range");
  if (hfx::launch_dft_axis(in, in_dtype, outer, n, inner, k_begin, k_end, scale,
                           reinterpret_cast<float2 *>(out), S(stream)) != 0)
    return fail(HFX_E_HIP, "dft launch failed");
  return HFX_OK;
}

static hfx_status fresnel_tables(hfx_fresnel *f, hipStream_t st) {
  hfx::launch_fresnel_chirp(f->lx, f->ly, f->H, f->W, f->lambda, f->z, f->dx, f->dy, st);
  HFX_TRY_HIP(hipGetLastError());
  return HFX_OK;
}

hfx_status hfx_fresnel_create(int64_t batch, int32_t height, int32_t width, double lambda_m,
                              double dx_m, double dy_m, double z_m, int32_t in_real,
                              int32_t output_magnitude, int32_t skip_phase_shift,
                              hfx_stream_t stream, hfx_fresnel **out) {
  if (!out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  if (batch <= 0 || height <= 0 || width <= 0) return fail(HFX_E_INVALID, "invalid shape");
  if (!(lambda_m > 0)) return fail(HFX_E_INVALID, "wavelength must be positive");
  if (!(dx_m > 0) || !(dy_m > 0)) return fail(HFX_E_INVALID, "pixel pitch must be positive");
  if (dx_m != dy_m) return fail(HFX_E_INVALID, "dx must equal dy");
  if (!(z_m != 0)) return fail(HFX_E_INVALID, "Fresnel z must be non-zero");
  auto *f = new (std::nothrow) hfx_fresnel();
  if (!f) return fail(HFX_E_NOMEM, "host allocation failed");
  std::unique_ptr<hfx_fresnel> guard(f);
  f->batch = batch;
  f->H = height;
  f->W = width;
  f->in_real = in_real;
  f->output_magnitude = output_magnitude;
  f->skip_phase_shift = skip_phase_shift;
  f->lambda = lambda_m;
  f->dx = dx_m;
  f->dy = dy_m;
  f->z = z_m;
  f->blk_c = hfx::pick_blk_c(width, height);
  if (!hfx::row_pass_supported(width) || !hfx::col_pass_supported(height, f->blk_c))
    return fail(HFX_E_UNSUPPORTED, "FFT size must be one of the compiled 2^a*3^b lengths (2..2048)");
  hipStream_t st = S(stream);
  HFX_TRY(f->arena.alloc(&f->tw_w, width));
  HFX_TRY(f->arena.alloc(&f->tw_h, height));
  HFX_TRY(f->arena.alloc(&f->lx, width));
  HFX_TRY(f->arena.alloc(&f->ly, height));
  HFX_TRY(f->arena.alloc(&f->ws, (size_t)batch * height * width));
  hfx::launch_fft_twiddles(f->tw_w, width, st);
  hfx::launch_fft_twiddles(f->tw_h, height, st);
  HFX_TRY(fresnel_tables(f, st));
  HFX_TRY_HIP(hipStreamSynchronize(st));
  *out = guard.release();
  return HFX_OK;
}

hfx_status hfx_fresnel_update_z(hfx_fresnel *f, double z_m, hfx_stream_t stream) {
  if (!f) return fail(HFX_E_INVALID, "null argument");
  if (!(z_m != 0)) return fail(HFX_E_INVALID, "Fresnel z must be non-zero");
  f->z = z_m;
  return fresnel_tables(f, S(stream));
}

hfx_status hfx_fresnel_execute(hfx_fresnel *f, const void *in, void *out, hfx_stream_t stream) {
  if (!f || !in || !out) return fail(HFX_E_INVALID, "null argument");
  hipStream_t st = S(stream);
  hfx::PassArgs r = base_pass(f->H, f->W, (int)f->batch, f->blk_c);
  r.in = in;
  r.in_mode = f->in_real ? hfx::IN_NAT_R : hfx::IN_NAT_C;
  r.pre_lx = f->lx;
  r.pre_ly = f->ly;
  r.out = f->ws;
  r.out_mode = hfx::OUT_BLK;
  r.tw = f->tw_w;
  if (hfx::launch_row_pass(r, st) != 0) return fail(HFX_E_HIP, "row pass launch failed");
  hfx::PassArgs c = base_pass(f->H, f->W, (int)f->batch, f->blk_c);
  c.in = f->ws;
  c.in_mode = hfx::IN_BLK;
  c.out = out;
  c.tw = f->tw_h;
  if (f->output_magnitude) {
    c.out_mode = hfx::OUT_NAT_MAG;
  } else {
    c.out_mode = hfx::OUT_NAT_C;
    if (!f->skip_phase_shift) {
      c.post_lx = f->lx;
      c.post_ly = f->ly;
    }
  }
  if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "column pass launch failed");
  return HFX_OK;
}

hfx_status hfx_fresnel_destroy(hfx_fresnel *f) {
  delete f;
  return HFX_OK;
}

hfx_status hfx_asm_create(const hfx_asm_desc *dp, hfx_stream_t stream, hfx_asm **out) {
  if (!dp || !out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  const hfx_asm_desc &d = *dp;
  if (d.batch <= 0 || d.height <= 0 || d.width <= 0) return fail(HFX_E_INVALID, "invalid shape");
  if (!(d.lambda_m > 0)) return fail(HFX_E_INVALID, "wavelength must be positive");
  if (!(d.dx_m > 0) || !(d.dy_m > 0)) return fail(HFX_E_INVALID, "pixel pitch must be positive");
  if (d.dx_m != d.dy_m) return fail(HFX_E_INVALID, "dx must equal dy");
  auto *a = new (std::nothrow) hfx_asm();
  if (!a) return fail(HFX_E_NOMEM, "host allocation failed");
  std::unique_ptr<hfx_asm> guard(a);
  a->d = d;
  a->OH = d.pad_height > 0 ? d.pad_height : d.height;
  a->OW = d.pad_width > 0 ? d.pad_width : d.width;
  if (a->OH < d.height || a->OW < d.width) return fail(HFX_E_INVALID, "padding smaller than input");
  if ((a->OH - d.height) % 2 || (a->OW - d.width) % 2)
    return fail(HFX_E_INVALID, "padding margins must be even");
  a->blk_c = hfx::pick_blk_c(a->OW, a->OH);
  if (!hfx::row_pass_supported(a->OW) || !hfx::col_pass_supported(a->OH, a->blk_c))
    return fail(HFX_E_UNSUPPORTED, "FFT size must be one of the compiled 2^a*3^b lengths (2..2048)");
  hipStream_t st = S(stream);
  HFX_TRY(a->arena.alloc(&a->tw_w, a->OW));
  HFX_TRY(a->arena.alloc(&a->tw_h, a->OH));
  HFX_TRY(a->arena.alloc(&a->tf, (size_t)a->OH * a->OW));
  HFX_TRY(a->arena.alloc(&a->ws, (size_t)d.batch * a->OH * a->OW));
  hfx::launch_fft_twiddles(a->tw_w, a->OW, st);
  hfx::launch_fft_twiddles(a->tw_h, a->OH, st);
  hfx::launch_asm_transfer(a->tf, a->OH, a->OW, d.lambda_m, d.z_m, d.dx_m, d.dy_m, d.filter_enabled,
                           d.r_inner, d.r_outer, d.s_inner, d.s_outer, a->blk_c, st);
  HFX_TRY_HIP(hipGetLastError());
  HFX_TRY_HIP(hipStreamSynchronize(st));
  *out = guard.release();
  return HFX_OK;
}

hfx_status hfx_asm_execute(hfx_asm *a, const void *in, void *out, hfx_stream_t stream) {
  if (!a || !in || !out) return fail(HFX_E_INVALID, "null argument");
  hipStream_t st = S(stream);
  const int B = (int)a->d.batch;
  hfx::PassArgs r = base_pass(a->OH, a->OW, B, a->blk_c);
  r.in = in;
  r.in_mode = a->d.in_real ? hfx::IN_NAT_R : hfx::IN_NAT_C;
  r.in_h = a->d.height;
  r.in_w = a->d.width;
  r.off_y = (a->OH - a->d.height) / 2;
  r.off_x = (a->OW - a->d.width) / 2;
  r.out = a->ws;
  r.out_mode = hfx::OUT_BLK;
  r.tw = a->tw_w;
  if (hfx::launch_row_pass(r, st) != 0) return fail(HFX_E_HIP, "row pass launch failed");
  hfx::PassArgs c = base_pass(a->OH, a->OW, B, a->blk_c);
  c.in = a->ws;
  c.in_mode = hfx::IN_BLK;
  c.mid_tf = a->tf;
  c.out = a->ws;
  c.out_mode = hfx::OUT_BLK;
  c.tw = a->tw_h;
  if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "column pass launch failed");
  hfx::PassArgs r2 = base_pass(a->OH, a->OW, B, a->blk_c);
  r2.in = a->ws;
  r2.in_mode = hfx::IN_BLK;
  r2.inv = 1;
  r2.out = out;
  r2.out_mode = hfx::OUT_NAT_C;
  r2.tw = a->tw_w;
  if (hfx::launch_row_pass(r2, st) != 0) return fail(HFX_E_HIP, "inverse row pass launch failed");
  return HFX_OK;
}

hfx_status hfx_asm_destroy(hfx_asm *a) {
  delete a;
  return HFX_OK;
}

}  // extern "C"

static hfx_status spectral_setup(SpectralFilter &f, int64_t batch, int H, int W, hipStream_t st,
                                 float2 *borrowed_ws) {
  if (batch <= 0 || H <= 0 || W <= 0) return fail(HFX_E_INVALID, "invalid shape");
  if (batch > 65535) return fail(HFX_E_UNSUPPORTED, "more than 65535 planes");
  f.batch = batch;
  f.H = H;
  f.W = W;
  f.blk_c = hfx::pick_blk_c(W, H);
  if (!hfx::row_pass_supported(W) || !hfx::col_pass_supported(H, f.blk_c))
    return fail(HFX_E_UNSUPPORTED, "FFT size must be one of the compiled 2^a*3^b lengths (2..2048)");
  HFX_TRY(f.arena.alloc(&f.tw_w, W));
  HFX_TRY(f.arena.alloc(&f.tw_h, H));
  HFX_TRY(f.arena.alloc(&f.tf, (size_t)H * W));
  if (borrowed_ws)
    f.ws = borrowed_ws;
  else
    HFX_TRY(f.arena.alloc(&f.ws, (size_t)batch * H * W));
  hfx::launch_fft_twiddles(f.tw_w, W, st);
  hfx::launch_fft_twiddles(f.tw_h, H, st);
  return HFX_OK;
}

static hfx_status spectral_run(const SpectralFilter &f, const void *in, bool in_real, void *out, bool out_real,
                               hipStream_t st) {
  const int B = (int)f.batch;
  hfx::PassArgs r = base_pass(f.H, f.W, B, f.blk_c);
  r.in = in;
  r.in_mode = in_real ? hfx::IN_NAT_R : hfx::IN_NAT_C;
  r.out = f.ws;
  r.out_mode = hfx::OUT_BLK;
  r.tw = f.tw_w;
  if (hfx::launch_row_pass(r, st) != 0) return fail(HFX_E_HIP, "row pass launch failed");
  hfx::PassArgs c = base_pass(f.H, f.W, B, f.blk_c);
  c.in = f.ws;
  c.in_mode = hfx::IN_BLK;
  c.mid_tf = f.tf;
  c.out = f.ws;
  c.out_mode = hfx::OUT_BLK;
  c.tw = f.tw_h;
  if (hfx::launch_col_pass(c, st) != 0) return fail(HFX_E_HIP, "column pass launch failed");
  hfx::PassArgs r2 = base_pass(f.H, f.W, B, f.blk_c);
  r2.in = f.ws;
  r2.in_mode = hfx::IN_BLK;
  r2.inv = 1;
  r2.out = out;
  r2.out_mode = out_real ? hfx::OUT_NAT_RE : hfx::OUT_NAT_C;
  r2.tw = f.tw_w;
  if (hfx::launch_row_pass(r2, st) != 0) return fail(HFX_E_HIP, "inverse row pass launch failed");
  return HFX_OK;
}

// Normalised Gaussian taps of radius ceil(3 sigma), in FP32 (flatfield.cu:73-94).
static bool gaussian_taps(float sigma, std::vector<float> &k) {
  const int radius = (int)std::ceil(3.0f * sigma);
  if (radius <= 0) return false;
  k.assign((size_t)(2 * radius + 1), 0.0f);
  const float inv_two_sigma2 = 1.0f / (2.0f * sigma * sigma);
  float sum = 0.0f;
  for (int i = -radius; i <= radius; ++i) {
    const float v = std::exp(-(float)(i * i) * inv_two_sigma2);
    k[(size_t)(i + radius)] = v;
    sum += v;
  }
  for (auto &v : k) v /= sum;
  return true;
}

// A float as the reference's generated callback source spells it: printf "%f" (six
// decimals), read back as a float literal (flatfield.cu:143-155).
static float six_decimals(float v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%f", (double)v);
  return std::strtof(buf, nullptr);
}

extern "C" {

hfx_status hfx_filter2d_create(int64_t batch, int32_t height, int32_t width, int32_t in_real, int32_t r_inner,
                               int32_t r_outer, int32_t s_inner, int32_t s_outer, hfx_stream_t stream,
                               hfx_filter2d **out) {
  if (!out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  if (r_inner < 0 || r_outer < 0 || s_inner < 0 || s_outer < 0)
    return fail(HFX_E_INVALID, "radii and tapers must be non-negative");
  auto f = std::make_unique<hfx_filter2d>();
  f->in_real = in_real ? 1 : 0;
  hipStream_t st = S(stream);
  HFX_TRY(spectral_setup(f->sf, batch, height, width, st));
  hfx::launch_filter2d_table(f->sf.tf, height, width, r_inner, r_outer, s_inner, s_outer, !f->in_real,
                             f->sf.blk_c, st);
  HFX_TRY_HIP(hipGetLastError());
  HFX_TRY_HIP(hipStreamSynchronize(st));
  *out = f.release();
  return HFX_OK;
}

hfx_status hfx_filter2d_execute(hfx_filter2d *f, const void *in, void *out, hfx_stream_t stream) {
  if (!f || !in || !out) return fail(HFX_E_INVALID, "null argument");
  return spectral_run(f->sf, in, f->in_real != 0, out, f->in_real != 0, S(stream));
}

hfx_status hfx_filter2d_destroy(hfx_filter2d *f) {
  delete f;
  return HFX_OK;
}

hfx_status hfx_flatfield_create(int64_t batch, int32_t height, int32_t width, float sigma_y, float sigma_x,
                                hfx_stream_t stream, hfx_flatfield **out, int32_t *uses_fft) {
  if (!out) return fail(HFX_E_INVALID, "null argument");
  *out = nullptr;
  if (batch <= 0 || height <= 0 || width <= 0) return fail(HFX_E_INVALID, "invalid shape");
  if (!(sigma_y > 0.0f) || !(sigma_x > 0.0f)) return fail(HFX_E_INVALID, "sigma must be positive");
  std::vector<float> ky, kx;
  if (!gaussian_taps(sigma_y, ky) || !gaussian_taps(sigma_x, kx))
    return fail(HFX_E_INVALID, "sigma is too small to form a Gaussian kernel");
  auto f = std::make_unique<hfx_flatfield>();
  f->batch = batch;
  f->H = height;
  f->W = width;
  f->ry = (int)(ky.size() / 2);
  f->rx = (int)(kx.size() / 2);
  // flatfield.cu:182-187: the Fourier form for kernels of 65 taps or more on 2-D planes
  f->use_fft = height > 1 && width > 1 && std::max(ky.size(), kx.size()) >= 65;
  hipStream_t st = S(stream);
  if (f->use_fft) {
    HFX_TRY(spectral_setup(f->sf, batch, height, width, st));
    const float scale = six_decimals(1.0f / (float)(width * height));
    hfx::launch_flatfield_table(f->sf.tf, height, width, six_decimals(sigma_y), six_decimals(sigma_x), scale,
                                f->sf.blk_c, st);
  } else {
    if (!hfx::flatfield_spatial_fused(f->rx, f->ry)) {
      HFX_TRY(f->arena.alloc(&f->tmp, (size_t)batch * height * width));
    } else if (batch > 65535) {
      return fail(HFX_E_UNSUPPORTED, "more than 65535 planes");
    }
    HFX_TRY(f->arena.alloc(&f->ky, ky.size()));
    HFX_TRY(f->arena.alloc(&f->kx, kx.size()));
    HFX_TRY_HIP(hipMemcpyAsync(f->ky, ky.data(), ky.size() * sizeof(float), hipMemcpyHostToDevice, st));
    HFX_TRY_HIP(hipMemcpyAsync(f->kx, kx.data(), kx.size() * sizeof(float), hipMemcpyHostToDevice, st));
  }
  HFX_TRY_HIP(hipGetLastError());
  HFX_TRY_HIP(hipStreamSynchronize(st));
  if (uses_fft) *uses_fft = f->use_fft ? 1 : 0;
  *out = f.release();
  return HFX_OK;
}

hfx_status hfx_flatfield_execute(hfx_flatfield *f, const float *in, float *out, hfx_stream_t stream) {
  if (!f || !in || !out) return fail(HFX_E_INVALID, "null argument");
  if (in == out) return fail(HFX_E_INVALID, "flatfield output must not alias its input");
  hipStream_t st = S(stream);
  if (f->use_fft) return spectral_run(f->sf, in, true, out, true, st);
  if (hfx::launch_flatfield_spatial(in, out, f->tmp, f->batch, f->H, f->W, f->kx, f->rx, f->ky, f->ry, st) != 0)
    return fail(HFX_E_HIP, "flatfield launch failed");
  return HFX_OK;
}

hfx_status hfx_flatfield_destroy(hfx_flatfield *f) {
  delete f;
  return HFX_OK;
}

hfx_status hfx_weighted_mean(const void *in, int32_t in_complex, int64_t planes, int64_t npix,
                             const float *weights, uint32_t flags, float *out, hfx_stream_t stream) {
  if (!in || !out) return fail(HFX_E_INVALID, "null argument");
  if (planes <= 0 || npix <= 0) return fail(HFX_E_INVALID, "reduction has zero elements");
  const int power2 = (flags & HFX_REDUCE_POWER2) ? 1 : 0;
  const int absval = (flags & HFX_REDUCE_ABS) ? 1 : 0;
  float *outs[3] = {out, nullptr, nullptr};
  int rc;
  if (weights) {
    // out = mean_k(w_k * m_k): the weighted sum is the reducer's second output; its first
    // (unweighted) output is written to the same buffer first and then overwritten.
    outs[1] = out;
    rc = hfx::launch_moments_reduce(in, in_complex, planes, npix, weights, weights, outs, 2, power2,
                                    absval, S(stream));
  } else {
    rc = hfx::launch_moments_reduce(in, in_complex, planes, npix, nullptr, nullptr, outs, 1, power2,
                                    absval, S(stream));
  }
  if (rc != 0) return fail(HFX_E_HIP, "reduce launch failed");
  return HFX_OK;
}

hfx_status hfx_strided_copy(const void *in, const int64_t *shape, const int64_t *in_strides_bytes,
                            int64_t in_offset_bytes, int32_t ndim, int32_t elem_size, void *out,
                            int64_t out_block, int64_t out_pitch, hfx_stream_t stream) {
  if (!in || !out || !shape || !in_strides_bytes) return fail(HFX_E_INVALID, "null argument");
  if (ndim < 1 || ndim > 8) return fail(HFX_E_INVALID, "rank must be in [1, 8]");
  if (elem_size <= 0 || in_offset_bytes < 0) return fail(HFX_E_INVALID, "invalid element size or offset");
  if (out_block < 0 || (out_block > 0 && out_pitch < out_block))
    return fail(HFX_E_INVALID, "out_pitch must be at least out_block");
  long long shp[8], st[8];
  for (int i = 0; i < ndim; ++i) {
    if (shape[i] < 0 || in_strides_bytes[i] < 0) return fail(HFX_E_INVALID, "negative shape or stride");
    shp[i] = shape[i];
    st[i] = in_strides_bytes[i];
  }
  if (hfx::launch_strided_copy(in, out, elem_size, ndim, shp, st, in_offset_bytes, out_block, out_pitch,
                               S(stream)) != 0)
    return fail(HFX_E_HIP, "copy launch failed");
  return HFX_OK;
}

hfx_status hfx_multiply(const void *a, int32_t a_complex, const int64_t *a_strides, const void *b,
                        int32_t b_complex, const int64_t *b_strides, int32_t ndim,
                        const int64_t *out_shape, void *out, hfx_stream_t stream) {
  if (!a || !b || !out || !a_strides || !b_strides || !out_shape) return fail(HFX_E_INVALID, "null argument");
  if (ndim < 1 || ndim > 8) return fail(HFX_E_INVALID, "rank must be in [1, 8]");
  long long shp[8], sa[8], sb[8];
  for (int i = 0; i < ndim; ++i) {
    if (out_shape[i] < 0 || a_strides[i] < 0 || b_strides[i] < 0)
      return fail(HFX_E_INVALID, "negative shape or stride");
    shp[i] = out_shape[i];
    sa[i] = a_strides[i];
    sb[i] = b_strides[i];
  }
  if (hfx::launch_multiply(a, a_complex, sa, b, b_complex, sb, ndim, shp, out, S(stream)) != 0)
    return fail(HFX_E_HIP, "multiply launch failed");
  return HFX_OK;
}

hfx_status hfx_fftshift(const void *in, void *out, int32_t elem_size, int32_t ndim,
                        const int64_t *shape, const int32_t *axes_mask, hfx_stream_t stream) {
  if (!in || !out || !shape || !axes_mask) return fail(HFX_E_INVALID, "null argument");
  if (in == out) return fail(HFX_E_INVALID, "fftshift is out of place");
  if (ndim < 1 || ndim > 8) return fail(HFX_E_INVALID, "rank must be in [1, 8]");
  if (elem_size <= 0) return fail(HFX_E_INVALID, "element size must be positive");
  long long shp[8];
  int msk[8];
  for (int i = 0; i < ndim; ++i) {
    if (shape[i] < 0) return fail(HFX_E_INVALID, "negative dimension");
    shp[i] = shape[i];
    msk[i] = axes_mask[i];
  }
  if (hfx::launch_fftshift(in, out, elem_size, ndim, shp, msk, S(stream)) != 0)
    return fail(HFX_E_HIP, "fftshift launch failed");
  return HFX_OK;
}

hfx_status hfx_convert_u16(const uint16_t *in, void *out, int64_t n, int32_t to_complex,
                           hfx_stream_t stream) {
  if ((!in || !out) && n > 0) return fail(HFX_E_INVALID, "null argument");
  if (hfx::launch_convert_u16(in, out, n, to_complex, S(stream)) != 0)
    return fail(HFX_E_HIP, "conversion launch failed");
  return HFX_OK;
}

int64_t hfx_convert_scratch_bytes(int64_t n) { return (int64_t)(hfx::minmax_blocks(n) + 1) * 8; }

hfx_status hfx_convert(const void *in, int32_t in_dtype, void *out, int32_t out_dtype, int32_t strategy,
                       int64_t n, void *scratch, hfx_stream_t stream) {
  if ((!in || !out) && n > 0) return fail(HFX_E_INVALID, "null argument");
  if (n < 0) return fail(HFX_E_INVALID, "negative element count");
  const bool ok =
      (strategy == HFX_CONV_REAL && (in_dtype == HFX_DTYPE_U8 || in_dtype == HFX_DTYPE_U16 || in_dtype == HFX_DTYPE_F32) &&
       (out_dtype == HFX_DTYPE_F32 || out_dtype == HFX_DTYPE_CF32) && !(in_dtype == HFX_DTYPE_F32 && out_dtype == HFX_DTYPE_F32)) ||
      (strategy == HFX_CONV_SCALED && in_dtype == HFX_DTYPE_F32 && (out_dtype == HFX_DTYPE_U8 || out_dtype == HFX_DTYPE_U16)) ||
      ((strategy == HFX_CONV_MODULUS || strategy == HFX_CONV_ARGUMENT) && in_dtype == HFX_DTYPE_CF32 && out_dtype == HFX_DTYPE_F32);
  if (!ok) return fail(HFX_E_INVALID, "unsupported conversion configuration");
  if (strategy == HFX_CONV_SCALED && !scratch && n > 0) return fail(HFX_E_INVALID, "Scaled needs a scratch buffer");
  if (hfx::launch_convert(in, in_dtype, out, out_dtype, strategy, n, scratch, S(stream)) != 0)
    return fail(HFX_E_HIP, "conversion launch failed");
  return HFX_OK;
}

hfx_status hfx_hbm_probe(int32_t mode, const void *in, void *out, int64_t bytes, int32_t blocks_per_cu,
                         int32_t unroll, int32_t nontemporal, hfx_stream_t stream) {
  if (mode < 0 || mode > 2) return fail(HFX_E_INVALID, "mode must be 0 (read), 1 (write) or 2 (copy)");
  if (bytes < 0 || blocks_per_cu < 1 || blocks_per_cu > 64) return fail(HFX_E_INVALID, "invalid size or grid");
  if (!out || (mode != 1 && !in)) return fail(HFX_E_INVALID, "null buffer");
  if ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) % 16) return fail(HFX_E_INVALID, "buffers must be 16-byte aligned");
  const int rc = hfx::launch_hbm_probe(mode, in, out, bytes, blocks_per_cu, unroll, nontemporal,
                                       reinterpret_cast<unsigned *>(out), S(stream));
  if (rc == -2) return fail(HFX_E_INVALID, "unroll must be 1, 2, 4 or 8");
  if (rc != 0) return fail(HFX_E_HIP, "probe launch failed");
  return HFX_OK;
}

int64_t hfx_convert_noalloc_scratch_bytes(void) { return hfx::convert_static_scratch_bytes(); }

hfx_status hfx_convert_noalloc(const void *in, int32_t in_dtype, void *out, int32_t out_dtype, int32_t strategy,
                               int64_t n, hfx_stream_t stream) {
  if (n < 0) return fail(HFX_E_INVALID, "negative element count");
  void *scratch = nullptr;
  if (strategy == HFX_CONV_SCALED && n > 0) {
    if (hfx_convert_scratch_bytes(n) > hfx::convert_static_scratch_bytes())
      return fail(HFX_E_INVALID, "hfx_convert_noalloc: " + std::to_string(n) + " elements need " +
                                     std::to_string(hfx_convert_scratch_bytes(n)) + " scratch bytes, more than the " +
                                     std::to_string(hfx::convert_static_scratch_bytes()) + " the library holds");
    if (hfx::convert_static_scratch(&scratch) != 0)
      return fail(HFX_E_HIP, "hfx_convert_noalloc: device address of the library scratch not available");
  }
  return hfx_convert(in, in_dtype, out, out_dtype, strategy, n, scratch, stream);
}

hfx_status hfx_slide_avg_step(const float *frame, float *old_slot, float *running, int64_t n, int32_t window,
                              hfx_stream_t stream) {
  if ((!frame || !old_slot || !running) && n > 0) return fail(HFX_E_INVALID, "null argument");
  if (window <= 0 || n < 0) return fail(HFX_E_INVALID, "window must be positive");
  if (hfx::launch_slide_avg(frame, old_slot, running, n, window, S(stream)) != 0)
    return fail(HFX_E_HIP, "sliding average launch failed");
  return HFX_OK;
}

void hfx_synth_desc_init(hfx_synth_desc *d, int32_t height, int32_t width) {
  std::memset(d, 0, sizeof(*d));
  d->height = height;
  d->width = width;
  d->disk_radius_px = 0.39 * (height < width ? height : width);
  d->fs_hz = 37e3;
  d->ref_amp = 1.0;
  d->obj_sigma = 0.5;
  d->gain = 1000.0;
  d->carrier_cyc_per_px = 0.25;
  d->phase_noise_rad = 0.3;
  d->fd_max_hz = 15e3;
  d->seed = 1234u;
  d->stream_id = 0u;
}

hfx_status hfx_synth_frames(const hfx_synth_desc *d, uint16_t *frames, int64_t first_frame,
                            int64_t n_frames, double *phase_state, hfx_stream_t stream) {
  if (!d || !frames || !phase_state) return fail(HFX_E_INVALID, "null argument");
  if (d->height <= 0 || d->width <= 0 || n_frames < 0) return fail(HFX_E_INVALID, "invalid shape");
  if (!(d->fs_hz > 0)) return fail(HFX_E_INVALID, "fs must be positive");
  hfx::SynthArgs a;
  a.H = d->height;
  a.W = d->width;
  a.disk_radius = d->disk_radius_px;
  a.fs = d->fs_hz;
  a.ref_amp = d->ref_amp;
  a.obj_sigma = d->obj_sigma;
  a.gain = d->gain;
  a.carrier = d->carrier_cyc_per_px;
  a.phase_noise = d->phase_noise_rad;
  a.fd_max = d->fd_max_hz;
  a.seed = d->seed;
  a.stream_id = d->stream_id;
  if (hfx::launch_synth(a, frames, first_frame, n_frames, phase_state, S(stream)) != 0)
    return fail(HFX_E_HIP, "synth launch failed");
  return HFX_OK;
}

}  // extern "C"
