// Stand-alone element-wise ops backing the per-node kinds of the path (the fused
// moment path does not use them):
//   fftshift  : out[i] = in[(i + n/2) mod n] on the selected axes (fftshift.cu:110-145)
//   moments   : mean over the plane axis of |x| (complex) or x (real magnitudes), with
//               optional f_k / f_k^2 weights (mean_abs.cu:239-276, mean.cu:268-306,
//               multiply.cu:69-92 wired by graph_builder.cc:612-631)
//   convert   : uint16 -> float32 or complex64 (x + 0i) (conversion.cu:156-202)
#include "hfx_kernels.hh"

namespace hfx {

struct ShiftGeo {
  int ndim;
  long long shape[8];
  long long stride[8];
  long long shift[8];
  long long total;
};

__global__ void k_fftshift(const uint8_t *in, uint8_t *out, int esz, ShiftGeo g) {
  long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= g.total) return;
  long long rem = tid, src = 0;
  for (int k = 0; k < g.ndim; ++k) {
    long long idx = g.stride[k] > 0 ? rem / g.stride[k] : 0;
    rem -= idx * g.stride[k];
    long long si = idx + g.shift[k];
    if (si >= g.shape[k]) si -= g.shape[k];
    src += si * g.stride[k];
  }
  if (esz == 4) {
    reinterpret_cast<uint32_t *>(out)[tid] = reinterpret_cast<const uint32_t *>(in)[src];
  } else if (esz == 8) {
    reinterpret_cast<uint2 *>(out)[tid] = reinterpret_cast<const uint2 *>(in)[src];
  } else {
    for (int b = 0; b < esz; ++b) out[tid * esz + b] = in[src * esz + b];
  }
}

int launch_fftshift(const void *in, void *out, int elem_size, int ndim, const long long *shape,
                    const int *mask, hipStream_t st) {
  if (ndim < 1 || ndim > 8) return -1;
  ShiftGeo g;
  g.ndim = ndim;
  long long acc = 1;
  for (int k = ndim - 1; k >= 0; --k) {
    g.shape[k] = shape[k];
    g.stride[k] = acc;
    g.shift[k] = mask[k] ? shape[k] / 2 : 0;
    acc *= shape[k];
  }
  g.total = acc;
  if (acc == 0) return 0;
  hipLaunchKernelGGL(k_fftshift, dim3((unsigned)((acc + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const uint8_t *>(in), reinterpret_cast<uint8_t *>(out),
                     elem_size, g);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

struct ReduceArgs {
  const void *in;
  int in_complex;
  long long planes, npix;
  const float *f, *f2;
  float *out[3];
  int n_out, power2;
};

__global__ void k_moments_reduce(ReduceArgs a) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= a.npix) return;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
  for (long long k = 0; k < a.planes; ++k) {
    float m;
    if (a.in_complex) {
      float2 v = reinterpret_cast<const float2 *>(a.in)[k * a.npix + p];
      m = v.x * v.x + v.y * v.y;
      if (!a.power2) m = sqrtf(m);
    } else {
      m = reinterpret_cast<const float *>(a.in)[k * a.npix + p];
      m = a.power2 ? m * m : fabsf(m);
    }
    s0 += m;
    if (a.n_out > 1) s1 += __fmul_rn(a.f[k], m);
    if (a.n_out > 2) s2 += __fmul_rn(a.f2[k], m);
  }
  const float n = (float)a.planes;
  a.out[0][p] = s0 / n;
  if (a.n_out > 1) a.out[1][p] = s1 / n;
  if (a.n_out > 2) a.out[2][p] = s2 / n;
}

int launch_moments_reduce(const void *in, int in_complex, long long planes, long long npix,
                          const float *f, const float *f2, float **out, int n_out, int power2,
                          hipStream_t st) {
  ReduceArgs a;
  a.in = in;
  a.in_complex = in_complex;
  a.planes = planes;
  a.npix = npix;
  a.f = f;
  a.f2 = f2;
  for (int i = 0; i < 3; ++i) a.out[i] = i < n_out ? out[i] : nullptr;
  a.n_out = n_out;
  a.power2 = power2;
  hipLaunchKernelGGL(k_moments_reduce, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

__global__ void k_convert_u16(const uint16_t *in, void *out, long long n, int to_complex) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float v = (float)in[i];
  if (to_complex)
    reinterpret_cast<float2 *>(out)[i] = make_float2(v, 0.f);
  else
    reinterpret_cast<float *>(out)[i] = v;
}

int launch_convert_u16(const uint16_t *in, void *out, long long n, int to_complex, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_convert_u16, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, out, n,
                     to_complex);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Direct DFT along the middle axis of [outer][n][inner] for the stand-alone RFFT / FFT
// kinds (any axis, any length).  Twiddles are evaluated in FP64 from the exactly
// reduced integer argument (k*j mod n) and rounded to FP32; products and sums are FP32.
struct DftAxisArgs {
  const void *in;
  int dtype;  // 1 U16, 2 F32, 3 CF32
  long long outer, inner;
  int n, kb, ke;
  float2 *out;
};

__global__ void k_dft_axis(DftAxisArgs a) {
  long long nk = a.ke - a.kb;
  long long total = a.outer * nk * a.inner;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  long long i = t % a.inner;
  long long k = (t / a.inner) % nk + a.kb;
  long long o = t / (a.inner * nk);
  float sr = 0.f, si = 0.f;
  for (int j = 0; j < a.n; ++j) {
    long long src = (o * a.n + j) * a.inner + i;
    float xr, xi = 0.f;
    if (a.dtype == 1)
      xr = (float)reinterpret_cast<const uint16_t *>(a.in)[src];
    else if (a.dtype == 2)
      xr = reinterpret_cast<const float *>(a.in)[src];
    else {
      float2 v = reinterpret_cast<const float2 *>(a.in)[src];
      xr = v.x;
      xi = v.y;
    }
    long long m = (k * j) % a.n;
    double s, c;
    sincospi(2.0 * (double)m / (double)a.n, &s, &c);
    float wr = (float)c, wi = (float)(-s);
    sr += xr * wr - xi * wi;
    si += xr * wi + xi * wr;
  }
  a.out[t] = make_float2(sr, si);
}

int launch_dft_axis(const void *in, int dtype, long long outer, int n, long long inner, int kb,
                    int ke, float2 *out, hipStream_t st) {
  DftAxisArgs a{in, dtype, outer, inner, n, kb, ke, out};
  long long total = outer * (long long)(ke - kb) * inner;
  if (total == 0) return 0;
  hipLaunchKernelGGL(k_dft_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hfx
