/*
 * hfx.h - C ABI of the MI355X (gfx950) Doppler-moment reconstruction path.
 *
 * Everything is plain C: device pointers, sizes and a stream handle (a hipStream_t passed
 * as void*).  No exception crosses this boundary: every entry point returns an hfx_status
 * and hfx_last_error() returns the thread-local message of the last failure.  The C++
 * holoflow node layer (include/holoflow/...) maps HFX_E_INVALID to std::invalid_argument
 * and HFX_E_HIP to curaii::HipError, as the reference's factories throw
 * std::invalid_argument from infer (fresnel_diffraction.cu:94-99) and curaii::CudaError
 * from the runtime (curaii/cuda.hh:29-36).
 *
 * Plans own their device tables and workspaces; they are allocated in *_create and freed
 * in *_destroy.  *_execute / hfx_moments* never allocate, never synchronise and only
 * enqueue work on the given stream, so they can be captured into a HIP graph.  A plan is
 * used by one thread / stream at a time; distinct plans may run concurrently.
 */
#ifndef HFX_H_
#define HFX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int hfx_status;
#define HFX_OK 0
#define HFX_E_INVALID 1     /* invalid argument (reference: std::invalid_argument) */
#define HFX_E_UNSUPPORTED 2 /* valid but outside what the kernels implement */
#define HFX_E_HIP 3         /* HIP runtime failure (reference: curaii::CudaError) */
#define HFX_E_NOMEM 4       /* device allocation failed */

typedef void *hfx_stream_t; /* a hipStream_t; NULL is the default stream */

/* Thread-local text of the last failure on this thread ("" if none). */
const char *hfx_last_error(void);
/* ABI version, bumped on any incompatible change. */
int32_t hfx_abi_version(void);

/* ------------------------------------------------------------------------------------
 * Fused Doppler-moment plan: uint16 window -> time DFT (band) -> Fresnel or angular-
 * spectrum propagation -> |X| -> M0/M1/M2 (+ band-split M0) -> optional fftshift.
 * Replaces the node chain GraphBuilder wires in build_time_frequency_analysis
 * (graph_builder.cc:245-305), build_spatial_propagation (:522-566) and build_xy_view
 * (:604-647): Conversion -> RFFT/FFT -> Slice/Copy/Concatenate -> FresnelDiffraction |
 * AngularSpectrum -> MeanAbs / Abs,Arange,Multiply,Mean -> FFTShift -> Mean(0).
 * ---------------------------------------------------------------------------------- */

enum hfx_prop { HFX_PROP_FRESNEL = 0, HFX_PROP_ANGULAR_SPECTRUM = 1 };
enum hfx_precision {
  HFX_PREC_FP32 = 0,      /* FP32 MFMA time DFT; every uint16 sample is exact */
  HFX_PREC_FP16 = 1,      /* FP16 MFMA, FP32 accumulate; samples exact up to 12 bits */
  HFX_PREC_BF16 = 2,      /* BF16 MFMA, FP32 accumulate; samples rounded to 8 significant bits */
  HFX_PREC_FP16_SPLIT = 3 /* FP16 MFMA with hi/lo split twiddles, FP32 accumulate; samples
                             exact up to 12 bits */
};
enum hfx_flags {
  HFX_FLAG_FFTSHIFT = 1u << 0,       /* fftshift the moment images (pp_fft_shift) */
  HFX_FLAG_POWER2 = 1u << 1,         /* integrate |X|^2 instead of the reference's |X| */
  HFX_FLAG_BOTH_SIDEBANDS = 1u << 2, /* TimeMethod::FFT: positive and negative band planes */
};

typedef struct hfx_plan_desc {
  int32_t height, width;    /* frame size H x W */
  int32_t window, stride;   /* STFT window T and stride S (frames) */
  double fs_hz;             /* camera frame rate (informational unless weight_fs_hz <= 0) */
  double lambda_m, dx_m, dy_m, z_m;
  int32_t bin_begin, bin_end; /* band bins [begin, end) of the T-point DFT, end <= T/2+1 */
  int32_t prop;             /* enum hfx_prop */
  int32_t precision;        /* enum hfx_precision */
  uint32_t flags;           /* enum hfx_flags */
  int32_t pad_height, pad_width; /* angular spectrum centred zero padding, 0 = none */
  int32_t filter_enabled;   /* annular filter (graph_builder.cc:124-126, 546-553): the angular
                               spectrum's transfer-function filter, or for Fresnel a Filter2D
                               stage on the band planes before propagation (integer radii) */
  double r_inner, r_outer, s_inner, s_outer;
  int32_t split_bin;        /* >0: also M0 over [begin, split) and [split, end) */
  double weight_fs_hz;      /* sampling rate of the M1/M2 weights f_k = k*fs/T; the
                               reference hard-codes 37000 (graph_builder.cc:580) */
  int32_t flatfield_enabled; /* Flatfield on every moment image after the fftshift
                                (graph_builder.cc:637-641) */
  float flatfield_sigma_y, flatfield_sigma_x;
  int32_t batch_windows;     /* hfx_moments_batch: windows per front launch (1..4; more keep
                                the pipelined front kernel busy across window boundaries at
                                the cost of one plane workspace each; FP32 Fresnel/ASM
                                without padding or Filter2D, else 1) */
  int32_t batch_lanes;       /* hfx_moments_batch: HIP streams the windows alternate over, so
                                that one window's kernels overlap the neighbouring window's
                                (0 or 2 = two unless a Flatfield or a Fresnel Filter2D stage
                                runs, else one; 1 = the caller's stream only).  Each lane
                                holds its own plane workspaces; the second lane runs on a
                                stream the plan owns and joins the caller's stream before the
                                call returns. */
} hfx_plan_desc;

typedef struct hfx_plan_info {
  int32_t out_height, out_width; /* moment image size (padded size for ASM padding) */
  int32_t n_outputs;             /* 3, or 5 with a band split */
  int32_t n_planes;              /* propagated planes (nb, or 2*nb for both sidebands) */
  int64_t workspace_bytes;       /* device bytes owned by the plan */
  int32_t batch_lanes;           /* streams hfx_moments_batch alternates over (1 or 2) */
  int32_t groupable;             /* 1: the plan's windows can join hfx_moments_group launches */
} hfx_plan_info;

typedef struct hfx_plan hfx_plan;

/* Fills a descriptor with the reference defaults (fs weights 37 kHz, FP32, Fresnel). */
void hfx_plan_desc_init(hfx_plan_desc *d);
/* Validates, allocates tables/workspaces and generates the FP64 phase tables on `stream`
 * (synchronised before return).  Replaces FresnelDiffractionFactory::create
 * (fresnel_diffraction.cu:510-632) / AngularSpectrumFactory::create
 * (angular_spectrum.cu:496-580) plus RFFTFactory::create (rfft.cu:264-297) for the path. */
hfx_status hfx_plan_create(const hfx_plan_desc *d, hfx_stream_t stream, hfx_plan **out);
hfx_status hfx_plan_destroy(hfx_plan *p);
hfx_status hfx_plan_get_info(const hfx_plan *p, hfx_plan_info *info);
/* Rewrites the propagation tables for a new z only (the reference's z-only reuse,
 * FresnelDiffractionFactory::update, fresnel_diffraction.cu:634-668). */
hfx_status hfx_plan_update_z(hfx_plan *p, double z_m, hfx_stream_t stream);
/* One window: frames first_frame .. first_frame+T-1 (mod ring_frames) of a device ring of
 * ring_frames frames of H*W uint16 each.  `out` is a device buffer of n_outputs*OH*OW
 * floats: M0, M1, M2 [, M0 low band, M0 high band].  Frame and pixel indices are 32-bit in
 * the kernels: rings of more than 2^31 - 1 frames (and, at hfx_plan_create, frames of more
 * than 2^31 - 1 pixels) return HFX_E_UNSUPPORTED. */
hfx_status hfx_moments(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                       int64_t first_frame, float *out, hfx_stream_t stream);
/* n_windows consecutive windows starting at first_frame, advancing by the plan stride;
 * window w writes out + w * out_stride floats. */
hfx_status hfx_moments_batch(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                             int64_t first_frame, int32_t n_windows, int64_t out_stride,
                             float *out, hfx_stream_t stream);
/* One window each of n plans (n <= HFX_GROUP_MAX, each plan at most once) in one front launch
 * and one launch per back pass: windows of different acquisitions batched together.  The
 * plans must agree on frame size, window, band bins, weights, propagation kind, precision
 * and flags (their z, wavelength and pitch may differ) and run the pipelined FP32 front
 * (frame width 512, 768 or 1024, up to 16 bins, T in {32, 64, 128}); HFX_E_INVALID names a
 * mismatch, HFX_E_UNSUPPORTED a geometry without the grouped kernels.  Window w reads
 * rings[w] (ring_frames[w] frames, from first_frames[w]) through plans[w]'s tables and
 * workspace and writes outs[w]; the results equal hfx_moments on each plan alone.  No
 * counterpart in the reference (one acquisition per process, manager.cc). */
#define HFX_GROUP_MAX 4
hfx_status hfx_moments_group(hfx_plan *const *plans, const uint16_t *const *rings, const int64_t *ring_frames,
                             const int64_t *first_frames, float *const *outs, int32_t n, hfx_stream_t stream);
/* Time-DFT stage of the plan alone (RFFT -> band slice, with the Fresnel lens applied
 * when the plan propagates with Fresnel): writes n_planes complex64 planes [n_planes][H][W]
 * to planes_out (device). */
hfx_status hfx_plan_time_dft(hfx_plan *p, const uint16_t *ring, int64_t ring_frames,
                             int64_t first_frame, void *planes_out, hfx_stream_t stream);
/* One half of hfx_moments, so that each kernel of the schedule can be timed alone:
 * stage 0 (front) = time DFT of the window and the forward row FFT (one fused kernel when
 * the width allows) into the plan's workspace; stage 1 (back) = the remaining passes and
 * the moment reduction from that workspace into out.  Stage 0 then stage 1 on one stream
 * is exactly hfx_moments(plan, ring, ring_frames, first_frame, out). */
hfx_status hfx_plan_run_stage(hfx_plan *p, int32_t stage, const uint16_t *ring, int64_t ring_frames,
                              int64_t first_frame, float *out, hfx_stream_t stream);

/* ------------------------------------------------------------------------------------
 * Per-node kinds (drop-in replacements of individual reference tasks).
 * ---------------------------------------------------------------------------------- */

/* holonp::RFFT / holonp::FFT along one axis (rfft.cu:148-297, fft.cu:234-386) for a
 * contiguous tensor viewed as [outer][n][inner] with the transform over n.  in_dtype:
 * 1 = U16, 2 = F32, 3 = CF32.  Outputs CF32 [outer][k_end-k_begin][inner] holding bins
 * [k_begin, k_end) times `scale` (1 for norm "backward", 1/n "forward", 1/sqrt(n) "ortho"). */
hfx_status hfx_dft_axis(const void *in, int32_t in_dtype, int64_t outer, int32_t n, int64_t inner,
                        int32_t k_begin, int32_t k_end, float scale, void *out, hfx_stream_t stream);

/* holotask::syncs::FresnelDiffraction (fresnel_diffraction.cu:386-668) over a contiguous
 * [batch][H][W] CF32 (in_real = 0) or F32 (in_real = 1) input, axes (-2, -1). */
typedef struct hfx_fresnel hfx_fresnel;
hfx_status hfx_fresnel_create(int64_t batch, int32_t height, int32_t width, double lambda_m,
                              double dx_m, double dy_m, double z_m, int32_t in_real,
                              int32_t output_magnitude, int32_t skip_phase_shift,
                              hfx_stream_t stream, hfx_fresnel **out);
hfx_status hfx_fresnel_update_z(hfx_fresnel *f, double z_m, hfx_stream_t stream);
hfx_status hfx_fresnel_execute(hfx_fresnel *f, const void *in, void *out, hfx_stream_t stream);
hfx_status hfx_fresnel_destroy(hfx_fresnel *f);

/* holotask::syncs::AngularSpectrum (angular_spectrum.cu:393-580): U = IFFT2_unnorm(
 * FFT2(pad(u)) * ifftshift(H)) over [batch][H][W] -> CF32 [batch][OH][OW]. */
typedef struct hfx_asm_desc {
  int64_t batch;
  int32_t height, width;
  double lambda_m, dx_m, dy_m, z_m;
  int32_t in_real;
  int32_t pad_height, pad_width; /* 0 = no padding */
  int32_t filter_enabled;
  double r_inner, r_outer, s_inner, s_outer;
} hfx_asm_desc;
typedef struct hfx_asm hfx_asm;
hfx_status hfx_asm_create(const hfx_asm_desc *d, hfx_stream_t stream, hfx_asm **out);
hfx_status hfx_asm_execute(hfx_asm *a, const void *in, void *out, hfx_stream_t stream);
hfx_status hfx_asm_destroy(hfx_asm *a);

/* holotask::syncs::Filter2D (filter2d.cu:165-497): out = IFFT2_unnorm(FFT2(in) * F) over
 * [batch][H][W], F the cosine-tapered annulus a(1-b) of radii r_inner/r_outer and tapers
 * s_inner/s_outer.  CF32 input (in_real = 0): F is the reference's centred table after the
 * corner exchange, a(1-b) in both real and imaginary parts, and `out` may alias `in`.
 * F32 input (in_real = 1): F is real on the DC-centred frequency grid and `out` is F32 (the
 * real-to-complex / complex-to-real pair).  Sizes: H and W in 2^a*3^b, 2..2048. */
typedef struct hfx_filter2d hfx_filter2d;
hfx_status hfx_filter2d_create(int64_t batch, int32_t height, int32_t width, int32_t in_real, int32_t r_inner,
                               int32_t r_outer, int32_t s_inner, int32_t s_outer, hfx_stream_t stream,
                               hfx_filter2d **out);
hfx_status hfx_filter2d_execute(hfx_filter2d *f, const void *in, void *out, hfx_stream_t stream);
hfx_status hfx_filter2d_destroy(hfx_filter2d *f);

/* holotask::syncs::Flatfield (flatfield.cu:73-451): out = in - GaussianBlur(in) per
 * [H][W] plane of F32 [batch][H][W].  Kernels of radius ceil(3 sigma) (sizes below 65, or a
 * one-pixel-wide or -tall plane): separable blur with clamp-at-edge indexing; larger
 * kernels: the Fourier-domain high-pass (1 - exp(-2 pi^2 (sx^2 fx^2 + sy^2 fy^2))) * scale,
 * with sigma and scale = 1/(H*W) rounded to six decimals as the reference's generated
 * source spells them.  `out` must not alias `in`.  *uses_fft reports the form chosen. */
typedef struct hfx_flatfield hfx_flatfield;
hfx_status hfx_flatfield_create(int64_t batch, int32_t height, int32_t width, float sigma_y, float sigma_x,
                                hfx_stream_t stream, hfx_flatfield **out, int32_t *uses_fft);
hfx_status hfx_flatfield_execute(hfx_flatfield *f, const float *in, float *out, hfx_stream_t stream);
hfx_status hfx_flatfield_destroy(hfx_flatfield *f);

/* holotask::syncs::MeanAbs (mean_abs.cu:231-276) / holonp::Mean (mean.cu:254-306) /
 * Multiply+Mean (multiply.cu:69-92) over the plane axis of [planes][npix]:
 *   out[p] = (1/planes) * sum_k w_k * m_k(p),  m = |x| for complex input, and x (or |x|
 * with HFX_REDUCE_ABS, as MeanAbs does) for real input; w_k = weights[k] (device array) or
 * 1 when weights is NULL.  HFX_REDUCE_POWER2: m = |x|^2. */
enum hfx_reduce_flags { HFX_REDUCE_POWER2 = 1u << 0, HFX_REDUCE_ABS = 1u << 1 };
hfx_status hfx_weighted_mean(const void *in, int32_t in_complex, int64_t planes, int64_t npix,
                             const float *weights, uint32_t flags, float *out, hfx_stream_t stream);

/* holonp::Copy (copy.cu:40,117-125): gathers the strided view (shape, byte strides, byte
 * offset from `in`) into `out`, element t of the view (row-major order) going to
 * out[(t / out_block) * out_pitch + t % out_block] (in elements; out_block = 0 writes
 * contiguously).  One call also places one input of holonp::Concatenate
 * (concatenate.cu:210,344-352) in its band of the output. */
hfx_status hfx_strided_copy(const void *in, const int64_t *shape, const int64_t *in_strides_bytes,
                            int64_t in_offset_bytes, int32_t ndim, int32_t elem_size, void *out,
                            int64_t out_block, int64_t out_pitch, hfx_stream_t stream);

/* holonp::Multiply with NumPy broadcasting (multiply.cu:69-92): out (contiguous, shape
 * out_shape) = a * b, operands given by element strides over out_shape (0 on broadcast
 * axes); a real operand times a complex one is promoted to x + 0i. */
hfx_status hfx_multiply(const void *a, int32_t a_complex, const int64_t *a_strides, const void *b,
                        int32_t b_complex, const int64_t *b_strides, int32_t ndim,
                        const int64_t *out_shape, void *out, hfx_stream_t stream);

/* holonp::FFTShift (fftshift.cu:110-145,254-260): out[i] = in[(i + n/2) mod n] on every
 * axis whose axes_mask entry is non-zero; elements of elem_size bytes. */
hfx_status hfx_fftshift(const void *in, void *out, int32_t elem_size, int32_t ndim,
                        const int64_t *shape, const int32_t *axes_mask, hfx_stream_t stream);

/* holotask::syncs::Conversion U16 -> F32 (real) or CF32 (x + 0i) (conversion.cu:156-202). */
hfx_status hfx_convert_u16(const uint16_t *in, void *out, int64_t n, int32_t to_complex,
                           hfx_stream_t stream);

/* holotask::syncs::Conversion, every pair the reference supports (conversion.cu:138-150,
 * 156-215, 299-340) plus U16 -> F32 Real (which the path's graph asks for,
 * graph_builder.cc:242).  Scaled maps the tensor's [min, max] to the full U8/U16 range,
 * int((x - min) * (top / (max - min)) + 0.5) clamped; it needs `scratch` of
 * hfx_convert_scratch_bytes(n) device bytes (per-block min/max partials). */
enum hfx_dtype { HFX_DTYPE_U16 = 1, HFX_DTYPE_F32 = 2, HFX_DTYPE_CF32 = 3, HFX_DTYPE_U8 = 4 };
enum hfx_conv_strategy { HFX_CONV_REAL = 0, HFX_CONV_SCALED = 1, HFX_CONV_MODULUS = 2, HFX_CONV_ARGUMENT = 3 };
int64_t hfx_convert_scratch_bytes(int64_t n);
hfx_status hfx_convert(const void *in, int32_t in_dtype, void *out, int32_t out_dtype, int32_t strategy,
                       int64_t n, void *scratch, hfx_stream_t stream);

/* hfx_convert without the scratch argument, for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The Scaled min/max partials live in a library-held
 * per-device array of hfx_convert_noalloc_scratch_bytes() bytes (8,200: enough for any n);
 * the call allocates nothing and rejects an n whose partials would not fit with
 * HFX_E_INVALID.  Calls on one GPU share that array: calls that overlap on one GPU (on
 * concurrent streams) race. */
int64_t hfx_convert_noalloc_scratch_bytes(void);
hfx_status hfx_convert_noalloc(const void *in, int32_t in_dtype, void *out, int32_t out_dtype, int32_t strategy,
                               int64_t n, hfx_stream_t stream);

/* One push of holotask::asyncs::SlidingAverage (slide_avg.cu:102-125, 200-215):
 * running = (running + frame / window) - old_slot / window, then old_slot = running. */
hfx_status hfx_slide_avg_step(const float *frame, float *old_slot, float *running, int64_t n, int32_t window,
                              hfx_stream_t stream);

/* HBM bandwidth probe (measurement only; no counterpart in the reference): mode 0 reads
 * `bytes` of `in`, 1 writes `bytes` of `out`, 2 copies `in` to `out`, with 16-byte vectors,
 * `unroll` (1, 2, 4, 8) independent vectors per lane in flight, blocks_per_cu x CUs
 * workgroups of 256 lanes, non-temporal accesses when `nontemporal` is set.  Buffers 16-byte
 * aligned; mode 0 uses the first KiB of `out` as a sink (its contents become undefined). */
hfx_status hfx_hbm_probe(int32_t mode, const void *in, void *out, int64_t bytes, int32_t blocks_per_cu,
                         int32_t unroll, int32_t nontemporal, hfx_stream_t stream);

/* ------------------------------------------------------------------------------------
 * Seeded synthetic interferograms (replaces the camera / holofile sources for tests and
 * benchmarks; no reference counterpart on the path).  phase_state: device array of
 * H*W doubles, zero before frame 0, carried between calls that continue a stream.
 * ---------------------------------------------------------------------------------- */
typedef struct hfx_synth_desc {
  int32_t height, width;
  double disk_radius_px;
  double fs_hz;
  double ref_amp, obj_sigma, gain, carrier_cyc_per_px, phase_noise_rad, fd_max_hz;
  uint32_t seed, stream_id;
} hfx_synth_desc;
void hfx_synth_desc_init(hfx_synth_desc *d, int32_t height, int32_t width);
hfx_status hfx_synth_frames(const hfx_synth_desc *d, uint16_t *frames, int64_t first_frame,
                            int64_t n_frames, double *phase_state, hfx_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* HFX_H_ */
