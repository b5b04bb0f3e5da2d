// Host-side launch interface of the gfx950 kernels.  Everything here is internal to
// libhfx; the public surface is the C ABI in include/hfx.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hfx_device.hh"

struct __half;
struct __hip_bfloat16;

namespace hfx {

// ---- FP64 table generation (tables.hip)
void launch_fft_twiddles(float2 *tw, int n, hipStream_t st);
void launch_dft_matrix(float *f32, __half *f16hi, __half *f16lo, __hip_bfloat16 *bf16, int rows,
                       int T, int kb, int nbc, hipStream_t st);
void launch_fresnel_chirp(float2 *lx, float2 *ly, int H, int W, double lambda, double z, double dx,
                          double dy, hipStream_t st);
void launch_asm_transfer(float2 *out, int H, int W, double lambda, double z, double dx, double dy,
                         int filt, double r_in, double r_out, double s_in, double s_out, int blk_c,
                         hipStream_t st);

// ---- optional in-path stages (filters.hip)
void launch_filter2d_table(float2 *out, int H, int W, int r_in, int r_out, int s_in, int s_out,
                           int complex_mode, int blk_c, hipStream_t st);
void launch_flatfield_table(float2 *out, int H, int W, float sigma_y, float sigma_x, float scale, int blk_c,
                            hipStream_t st);
bool flatfield_spatial_fused(int rx, int ry);
int launch_flatfield_spatial(const float *in, float *out, float *tmp, long long planes, int H, int W,
                             const float *kx, int rx, const float *ky, int ry, hipStream_t st);

// ---- time DFT (dft.hip)
constexpr int HFX_MAX_WIN = 4;  // windows per pipelined front launch (one plan, or a group of plans)

struct DftLaunch {
  const uint16_t *ring;
  long long ring_frames, first;
  // pipelined front only: windows 1 .. nwin-1 start at firsts[w] and write their planes at
  // planes + w * planes_stride (nwin = 1: just `first` and `planes`)
  int nwin;
  long long firsts[HFX_MAX_WIN];
  long long planes_stride;
  // grouped launch (windows of several plans of one shape): per-window ring, ring length,
  // planes and Fresnel lens; null rings[w] means ring / ring_frames / planes + w * stride / lx, ly
  const uint16_t *rings[HFX_MAX_WIN];
  long long ring_frames_w[HFX_MAX_WIN];
  float2 *planes_w[HFX_MAX_WIN];
  const float2 *lx_w[HFX_MAX_WIN], *ly_w[HFX_MAX_WIN];
  int T;
  long long npix;
  int W;
  int kb, nb, nbc, both;
  float2 *planes;
  const float2 *lx, *ly;
  int precision;  // 0 FP32, 1 FP16, 2 BF16, 3 FP16 split
  const float *wm_f32;
  const uint16_t *wm_f16hi, *wm_f16lo, *wm_bf16;
};
int launch_time_dft(const DftLaunch &L, hipStream_t st);
// Fused time DFT + forward row FFT into the column-blocked layout (L.planes is blocked).
bool dft_row_supported(int W, int nbc, int T);
// true when launch_time_dft_row runs the pipelined front kernel (and so accepts nwin > 1)
bool dft_front_pipelined(int W, int nbc, int T, int n_planes, int precision);
int launch_time_dft_row(const DftLaunch &L, int H, int blk_c, int n_planes, const float2 *tw,
                        hipStream_t st);

// ---- 2-D FFT passes (fft2d.hip)
enum InMode { IN_NAT_C = 0, IN_NAT_R = 1, IN_BLK = 2 };
enum OutMode { OUT_BLK = 0, OUT_NAT_C = 1, OUT_NAT_MAG = 2, OUT_MOMENTS = 3, OUT_NAT_RE = 4 };

struct MomentArgs {
  float *out[5];       // M0, M1, M2, M0 low band, M0 high band (out[3..4] may be null)
  const float *f;      // [planes] weights f_k (FP32, arange semantics)
  const float *f2;     // [planes] f_k * f_k (FP32)
  int split;           // band split plane index (0 = none)
  int power2;          // accumulate |X|^2 instead of |X|
  int shift;           // fftshift the stored images on both axes
};

// One window of a grouped pass launch (windows of several plans of one shape, blockIdx.z):
// the pointers that differ between the windows; everything else is shared.
struct PassWin {
  const void *in;
  void *out;
  const float2 *mid_tf;
  float *mom_out[5];
};

struct PassArgs {
  int H, W;                 // transform grid (output plane size)
  int planes;               // number of planes
  int inv;                  // inverse transform (exp(+i...)), unnormalised
  int blk_c;                // column-block width of the blocked layout
  // input
  const void *in;
  int in_mode;
  int in_h, in_w;           // input plane dims (== H, W unless zero-padded)
  int off_y, off_x;         // centred padding offsets
  const float2 *pre_lx, *pre_ly;  // lens applied to the input (row pass) or null
  // column pass: multiply by a blocked table between forward and inverse column FFTs
  const float2 *mid_tf;
  // output
  void *out;
  int out_mode;
  const float2 *post_lx, *post_ly;  // lens applied to the output or null
  MomentArgs mom;
  const float2 *tw;         // FFT twiddles for the transform length
  const float2 *tw2;        // second twiddle table unused by single-axis passes
  int nwin;                 // > 1: grouped launch, window blockIdx.z takes win[blockIdx.z]
  PassWin win[HFX_MAX_WIN];
};

int launch_row_pass(const PassArgs &a, hipStream_t st);
int launch_col_pass(const PassArgs &a, hipStream_t st);
bool row_pass_supported(int W);
bool col_pass_supported(int H, int blk_c);
int pick_blk_c(int W, int H);
// wave-level FFT passes (wpass.hip) for lengths 256, 512, 768, 1024, 2048
bool wave_pass_supported(int n);
int launch_wave_row_pass(const PassArgs &a, hipStream_t st);
int launch_wave_col_pass(const PassArgs &a, hipStream_t st);

// ---- small ops (ops.hip)
int launch_fftshift(const void *in, void *out, int elem_size, int ndim, const long long *shape,
                    const int *shift_axes_mask, hipStream_t st);
int launch_moments_reduce(const void *in, int in_complex, long long planes, long long npix,
                          const float *f, const float *f2, float **out, int n_out, int power2,
                          int absval, hipStream_t st);
int launch_strided_copy(const void *in, void *out, int elem_size, int ndim, const long long *shape,
                        const long long *in_strides, long long in_offset, long long out_block,
                        long long out_pitch, hipStream_t st);
int launch_multiply(const void *a, int a_complex, const long long *sa, const void *b, int b_complex,
                    const long long *sb, int ndim, const long long *shape, void *out, hipStream_t st);
int launch_convert_u16(const uint16_t *in, void *out, long long n, int to_complex, hipStream_t st);
int minmax_blocks(long long n);
long long convert_static_scratch_bytes();
int convert_static_scratch(void **ptr);
int launch_convert(const void *in, int in_dt, void *out, int out_dt, int strategy, long long n,
                   void *scratch, hipStream_t st);
int launch_hbm_probe(int mode, const void *in, void *out, long long bytes, int blocks_per_cu, int unroll,
                     int nontemporal, unsigned *sink, hipStream_t st);
int launch_slide_avg(const float *x, float *old_slot, float *ravg, long long n, int ws, hipStream_t st);
int launch_dft_axis(const void *in, int dtype, long long outer, int n, long long inner, int kb,
                    int ke, float scale, float2 *out, hipStream_t st);

// ---- synthetic interferograms (synth.hip)
struct SynthArgs {
  int H, W;
  double disk_radius;
  double fs;
  double ref_amp, obj_sigma, gain;
  double carrier;
  double phase_noise;
  double fd_max;
  uint32_t seed, stream_id;
};
int launch_synth(const SynthArgs &a, uint16_t *frames, long long first_frame, long long n_frames,
                 double *phase_state, hipStream_t st);

}  // namespace hfx
