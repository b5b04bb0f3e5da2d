// FP64 generation of every phase table on the path.  Phases are evaluated in double on
// the GPU and rounded once to FP32 (or FP16/BF16 hi/lo pairs for the MFMA operands),
// because the propagation phases reach 1e6 rad where one FP32 ulp is ~0.1 rad.
#include "hfx_kernels.hh"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace hfx {

// tw[m] = exp(-2*pi*i*m/N), m in [0, N).
__global__ void k_fft_twiddles(float2 *tw, int n) {
  int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  double s, c;
  sincospi(2.0 * (double)m / (double)n, &s, &c);
  tw[m] = make_float2((float)c, (float)(-s));
}

void launch_fft_twiddles(float2 *tw, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_fft_twiddles, dim3((n + 255) / 256), dim3(256), 0, st, tw, n);
}

// Time-DFT operand matrix: rows 2b / 2b+1 hold cos / -sin of 2*pi*(kb+b)*n/T for the
// band bins b in [0, nbc); padded rows (b >= nbc) are zero.  The argument is reduced
// exactly in integers ((kb+b)*n mod T) before the FP64 sincospi.
//   f32   : [rows][T] float
//   f16hi : [rows][T] half    (round(w))
//   f16lo : [rows][T] half    (round(w - hi))
//   bf16  : [rows][T] bf16
__global__ void k_dft_matrix(float *f32, __half *f16hi, __half *f16lo, __hip_bfloat16 *bf16,
                             int rows, int T, int kb, int nbc) {
  int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= rows * T) return;
  int row = idx / T, n = idx % T;
  int b = row >> 1;
  double w = 0.0;
  if (b < nbc) {
    long long m = ((long long)(kb + b) * (long long)n) % T;
    double s, c;
    sincospi(2.0 * (double)m / (double)T, &s, &c);
    w = (row & 1) ? -s : c;
  }
  f32[idx] = (float)w;
  // double -> binary16 / bfloat16 with a single round-to-nearest-even
  _Float16 hi = (_Float16)w;
  _Float16 lo = (_Float16)(w - (double)hi);
  __bf16 wb = (__bf16)w;
  reinterpret_cast<_Float16 *>(f16hi)[idx] = hi;
  reinterpret_cast<_Float16 *>(f16lo)[idx] = lo;
  reinterpret_cast<__bf16 *>(bf16)[idx] = wb;
}

void launch_dft_matrix(float *f32, __half *f16hi, __half *f16lo, __hip_bfloat16 *bf16, int rows,
                       int T, int kb, int nbc, hipStream_t st) {
  int n = rows * T;
  hipLaunchKernelGGL(k_dft_matrix, dim3((n + 255) / 256), dim3(256), 0, st, f32, f16hi, f16lo,
                     bf16, rows, T, kb, nbc);
}

// Fresnel quadratic lens, separable form.  The reference lens is
//   L[r,c] = exp(i*pi/(lambda*z) * (x^2 + y^2)),
//   x = ((c + (S-W)/2) - S/2) * dx,  y = ((r + (S-H)/2) - S/2) * dy,  S = max(H, W)
// (fresnel_diffraction.cu:326-341; S/2 is a real division there).  Since
// exp(i*a*(x^2+y^2)) = exp(i*a*x^2) * exp(i*a*y^2), two vectors carry it exactly:
//   lx[c] = exp(i*a*x_c^2), ly[r] = exp(i*a*y_r^2).
__global__ void k_fresnel_chirp(float2 *lx, float2 *ly, int H, int W, double lambda, double z,
                                double dx, double dy) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int S = H > W ? H : W;
  double a = M_PI / (lambda * z);
  if (i < W) {
    int off = (S - W) / 2;
    double x = ((double)(i + off) - (double)S / 2.0) * dx;
    double s, c;
    sincos(a * x * x, &s, &c);
    lx[i] = make_float2((float)c, (float)s);
  }
  if (i < H) {
    int off = (S - H) / 2;
    double y = ((double)(i + off) - (double)S / 2.0) * dy;
    double s, c;
    sincos(a * y * y, &s, &c);
    ly[i] = make_float2((float)c, (float)s);
  }
}

void launch_fresnel_chirp(float2 *lx, float2 *ly, int H, int W, double lambda, double z, double dx,
                          double dy, hipStream_t st) {
  int n = H > W ? H : W;
  hipLaunchKernelGGL(k_fresnel_chirp, dim3((n + 255) / 256), dim3(256), 0, st, lx, ly, H, W,
                     lambda, z, dx, dy);
}

// Angular-spectrum transfer function (angular_spectrum.cu:261-387), generated in FP64:
//   centred grid: u = (c - W/2)/(W*dx), v = (r - H/2)/(H*dy)  (integer W/2, H/2)
//   Hc = exp(i * 2*pi*z/lambda * sqrt(max(0, 1 - lambda^2 (u^2+v^2))))
//   optional annular filter: Hc *= w(dist)^2 with dist from (W/2.0, H/2.0)
//   stored after the reference's swap_corners quadrant exchange (ifftshift for even sizes).
// `blk_c` > 0 stores the table in the column-blocked layout used by the fused kernels:
//   (r, c) -> (c / blk_c) * H * blk_c + r * blk_c + c % blk_c.
__global__ void k_asm_transfer(float2 *out, int H, int W, double lambda, double z, double dx,
                               double dy, int filt, double r_in, double r_out, double s_in,
                               double s_out, int blk_c) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  int r = blockIdx.y;
  if (c >= W || r >= H) return;
  int hw = W / 2, hh = H / 2;
  int rs = r, cs = c;
  if (r < 2 * hh && c < 2 * hw) {
    rs = r < hh ? r + hh : r - hh;
    cs = c < hw ? c + hw : c - hw;
  }
  double du = 1.0 / ((double)W * dx);
  double dv = 1.0 / ((double)H * dy);
  double u = (double)(cs - W / 2) * du;
  double v = (double)(rs - H / 2) * dv;
  double tmp = 1.0 - lambda * lambda * (u * u + v * v);
  tmp = tmp > 0.0 ? tmp : 0.0;
  double phase = 2.0 * M_PI * z / lambda * sqrt(tmp);
  double s, co;
  sincos(phase, &s, &co);
  double wgt = 1.0;
  if (filt) {
    double rx = (double)cs - (double)W / 2.0;
    double ry = (double)rs - (double)H / 2.0;
    double dist = hypot(rx, ry);
    double inner_start = r_in - s_in, inner_end = r_in + s_in;
    double outer_start = r_out - s_out, outer_end = r_out + s_out;
    double val;
    if (dist < inner_start)
      val = 0.0;
    else if (dist < inner_end)
      val = 0.5 * (1.0 + sin(M_PI / (2.0 * s_in) * (dist - inner_start)));
    else if (dist < outer_start)
      val = 1.0;
    else if (dist < outer_end)
      val = 0.5 * (1.0 + sin(M_PI / (2.0 * s_out) * (outer_end - dist)));
    else
      val = 0.0;
    wgt = val * val;
  }
  float2 h = make_float2((float)(co * wgt), (float)(s * wgt));
  long long o;
  if (blk_c > 0)
    o = (long long)(c / blk_c) * H * blk_c + (long long)r * blk_c + (c % blk_c);
  else
    o = (long long)r * W + c;
  out[o] = h;
}

void launch_asm_transfer(float2 *out, int H, int W, double lambda, double z, double dx, double dy,
                         int filt, double r_in, double r_out, double s_in, double s_out, int blk_c,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_asm_transfer, dim3((W + 127) / 128, H), dim3(128), 0, st, out, H, W, lambda,
                     z, dx, dy, filt, r_in, r_out, s_in, s_out, blk_c);
}

}  // namespace hfx
