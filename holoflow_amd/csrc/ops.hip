// Stand-alone element-wise ops backing the per-node kinds of the path (the fused
// moment path does not use them):
//   fftshift  : out[i] = in[(i + n/2) mod n] on the selected axes (fftshift.cu:110-145)
//   moments   : mean over the plane axis of |x| (complex) or x (real magnitudes), with
//               optional f_k / f_k^2 weights (mean_abs.cu:239-276, mean.cu:268-306,
//               multiply.cu:69-92 wired by graph_builder.cc:612-631)
//   convert   : uint16 -> float32 or complex64 (x + 0i) (conversion.cu:156-202)
#include <mutex>

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
  int n_out, power2, absval;
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
      m = a.power2 ? m * m : (a.absval ? fabsf(m) : m);
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
                          int absval, hipStream_t st) {
  ReduceArgs a;
  a.absval = absval;
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

// ---------------------------------------------------------------- Conversion (all pairs)
// dtype codes: 1 U16, 2 F32, 3 CF32, 4 U8; strategy: 0 Real, 1 Scaled, 2 Modulus,
// 3 Argument (conversion.cu:138-150, kernels :156-215).
__device__ __forceinline__ float conv_load_real(const void *in, int dt, long long i) {
  switch (dt) {
    case 1: return (float)reinterpret_cast<const uint16_t *>(in)[i];
    case 4: return (float)reinterpret_cast<const uint8_t *>(in)[i];
    default: return reinterpret_cast<const float *>(in)[i];
  }
}

__global__ void k_convert(const void *in, int in_dt, void *out, int out_dt, int strategy, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (in_dt == 3) {  // CF32 -> F32: modulus (hypot) or argument (atan2)
    const float2 z = reinterpret_cast<const float2 *>(in)[i];
    reinterpret_cast<float *>(out)[i] = strategy == 3 ? atan2f(z.y, z.x) : hypotf(z.x, z.y);
    return;
  }
  const float v = conv_load_real(in, in_dt, i);
  if (out_dt == 3)
    reinterpret_cast<float2 *>(out)[i] = make_float2(v, 0.f);
  else
    reinterpret_cast<float *>(out)[i] = v;
}

// min / max of n floats: per-block partials, then one block folds them (order-free ops)
__global__ __launch_bounds__(256) void k_minmax_partial(const float *in, long long n, float2 *partial) {
  float lo = INFINITY, hi = -INFINITY;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    const float v = in[i];
    lo = fminf(lo, v);
    hi = fmaxf(hi, v);
  }
  __shared__ float2 red[256];
  red[threadIdx.x] = make_float2(lo, hi);
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x].x = fminf(red[threadIdx.x].x, red[threadIdx.x + s].x);
      red[threadIdx.x].y = fmaxf(red[threadIdx.x].y, red[threadIdx.x + s].y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_minmax_final(float2 *partial, int nblocks) {
  float lo = INFINITY, hi = -INFINITY;
  for (int i = threadIdx.x; i < nblocks; i += 256) {
    lo = fminf(lo, partial[i].x);
    hi = fmaxf(hi, partial[i].y);
  }
  __shared__ float2 red[256];
  red[threadIdx.x] = make_float2(lo, hi);
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[threadIdx.x].x = fminf(red[threadIdx.x].x, red[threadIdx.x + s].x);
      red[threadIdx.x].y = fmaxf(red[threadIdx.x].y, red[threadIdx.x + s].y);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[nblocks] = red[0];
}

// F32 -> U8/U16 scaled to the tensor's [min, max] (conversion.cu:177-195)
__global__ void k_scaled(const float *in, void *out, int out_dt, long long n, const float2 *mm) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float lo = mm->x, hi = mm->y;
  const float top = out_dt == 4 ? 255.f : 65535.f;
  const float scale = top / (hi - lo);
  const int val = (int)((in[i] - lo) * scale + 0.5f);
  const float c = fminf(fmaxf((float)val, 0.f), top);
  if (out_dt == 4)
    reinterpret_cast<uint8_t *>(out)[i] = (uint8_t)c;
  else
    reinterpret_cast<uint16_t *>(out)[i] = (uint16_t)c;
}

int minmax_blocks(long long n) { return (int)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024); }

// Library-held scratch of the allocation-free Scaled conversion: the per-block min/max
// partials of at most 1024 blocks plus the folded pair (8,200 bytes, any n).  Each GPU holds
// its own copy for the life of the process; calls on one GPU share it, so calls that overlap
// on one GPU (concurrent streams) race.  Every call writes each pair it reads first, so
// nothing needs clearing between calls.
constexpr int kConvScratchPairs = 1025;
static __device__ __align__(256) float2 scratch[kConvScratchPairs];

long long convert_static_scratch_bytes() { return (long long)sizeof(float2) * kConvScratchPairs; }

// Device address of `scratch` on the current device, looked up once per device (the host
// symbol is only a stand-in: kernels must receive this address).
int convert_static_scratch(void **ptr) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static void *per_device[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return -1;
  std::lock_guard<std::mutex> lock(mu);
  if (!per_device[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return -1;
    per_device[dev] = p;
  }
  *ptr = per_device[dev];
  return 0;
}

int launch_convert(const void *in, int in_dt, void *out, int out_dt, int strategy, long long n,
                   void *scratch, hipStream_t st) {
  if (n == 0) return 0;
  const unsigned grid = (unsigned)((n + 255) / 256);
  if (strategy == 1) {
    const int nb = minmax_blocks(n);
    float2 *part = reinterpret_cast<float2 *>(scratch);
    hipLaunchKernelGGL(k_minmax_partial, dim3(nb), dim3(256), 0, st, reinterpret_cast<const float *>(in), n, part);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, st, part, nb);
    hipLaunchKernelGGL(k_scaled, dim3(grid), dim3(256), 0, st, reinterpret_cast<const float *>(in), out, out_dt,
                       n, part + nb);
  } else {
    hipLaunchKernelGGL(k_convert, dim3(grid), dim3(256), 0, st, in, in_dt, out, out_dt, strategy, n);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- SlidingAverage step
// ravg = (ravg + x / ws) - old / ws, then the old slot takes ravg (slide_avg.cu:102-125,
// 200-215: two kernels and a copy; each operation rounds to FP32 in the same order here).
__global__ void k_slide_avg(const float *x, float *old_slot, float *ravg, long long n, float ws) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float r = ravg[i];
  r = r + x[i] / ws;
  r = r - old_slot[i] / ws;
  ravg[i] = r;
  old_slot[i] = r;
}

int launch_slide_avg(const float *x, float *old_slot, float *ravg, long long n, int ws, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_slide_avg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, old_slot, ravg, n,
                     (float)ws);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Gather of a strided view (holonp::Copy, copy.cu:40,117-125; the pieces of
// holonp::Concatenate).  Strides and offset in bytes, as TDesc has them.  Element t of the
// view goes to out[(t / out_block) * out_pitch + t % out_block]: one input of a
// concatenation along an inner axis lands in its column band of the output in one launch.
struct CopyGeo {
  int ndim, esz;
  long long shape[8];
  long long in_stride[8];
  long long in_offset;
  long long total;
  long long out_block, out_pitch;
};

__global__ void k_strided_copy(const uint8_t *in, uint8_t *out, CopyGeo g) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.total) return;
  long long rem = t, src = g.in_offset;
  for (int k = g.ndim - 1; k >= 0; --k) {
    const long long idx = rem % g.shape[k];
    rem /= g.shape[k];
    src += idx * g.in_stride[k];
  }
  const long long o = (t / g.out_block) * g.out_pitch + t % g.out_block;
  if (g.esz == 8)
    reinterpret_cast<uint2 *>(out)[o] = *reinterpret_cast<const uint2 *>(in + src);
  else if (g.esz == 4)
    reinterpret_cast<uint32_t *>(out)[o] = *reinterpret_cast<const uint32_t *>(in + src);
  else
    for (int b = 0; b < g.esz; ++b) out[o * g.esz + b] = in[src + b];
}

// Contiguous fast path: 16 bytes per lane, 4 vectors in flight per lane, grid sized to the
// device (8 blocks of 256 per CU), grid-stride over the buffer.
__global__ __launch_bounds__(256) void k_copy16(const uint4 *__restrict__ in, uint4 *__restrict__ out,
                                                long long n16) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n16; i += 4 * stride) {
    const uint4 a = in[i], b = in[i + stride], c = in[i + 2 * stride], d = in[i + 3 * stride];
    out[i] = a;
    out[i + stride] = b;
    out[i + 2 * stride] = c;
    out[i + 3 * stride] = d;
  }
  for (; i < n16; i += stride) out[i] = in[i];
}

// HBM bandwidth probe (measurement only): mode 0 reads the buffer (each lane folds its
// 16-byte vectors into one word, stored only if it equals a value no data here produces, so
// the loads stay live), mode 1 writes it, mode 2 copies it.  U independent 16-byte vectors per
// lane per iteration, grid-stride over the buffer, non-temporal for NT = 1.
template <int MODE, int U, int NT>
__global__ __launch_bounds__(256) void k_hbm_probe(const uint4 *__restrict__ in, uint4 *__restrict__ out,
                                                   long long n16, unsigned *sink) {
  const long long stride = (long long)gridDim.x * 256;
  long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  typedef unsigned u4v __attribute__((ext_vector_type(4)));
  unsigned acc = 0;
  const uint4 fill = make_uint4(threadIdx.x, blockIdx.x, 0x5a5a5a5au, 0xa5a5a5a5u);
  for (; i + (U - 1) * stride < n16; i += U * stride) {
    u4v v[U];
    if (MODE != 1) {
#pragma unroll
      for (int u = 0; u < U; ++u)
        v[u] = NT ? __builtin_nontemporal_load(reinterpret_cast<const u4v *>(in) + i + u * stride)
                  : reinterpret_cast<const u4v *>(in)[i + u * stride];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if (MODE == 0) {
        acc ^= v[u].x ^ v[u].y ^ v[u].z ^ v[u].w;
      } else {
        const u4v w = MODE == 1 ? u4v{fill.x, fill.y, fill.z, fill.w} : v[u];
        if (NT)
          __builtin_nontemporal_store(w, reinterpret_cast<u4v *>(out) + i + u * stride);
        else
          reinterpret_cast<u4v *>(out)[i + u * stride] = w;
      }
    }
  }
  for (; i < n16; i += stride) {
    if (MODE == 0) {
      const uint4 v = in[i];
      acc ^= v.x ^ v.y ^ v.z ^ v.w;
    } else {
      out[i] = MODE == 1 ? fill : in[i];
    }
  }
  if (MODE == 0 && acc == 0x9e3779b9u) sink[blockIdx.x & 255] = acc;
}

template <int MODE, int U>
static void launch_probe_u(int nt, unsigned blocks, const uint4 *in, uint4 *out, long long n16, unsigned *sink,
                           hipStream_t st) {
  if (nt)
    hipLaunchKernelGGL((k_hbm_probe<MODE, U, 1>), dim3(blocks), dim3(256), 0, st, in, out, n16, sink);
  else
    hipLaunchKernelGGL((k_hbm_probe<MODE, U, 0>), dim3(blocks), dim3(256), 0, st, in, out, n16, sink);
}

template <int MODE>
static int launch_probe_mode(int unroll, int nt, unsigned blocks, const uint4 *in, uint4 *out, long long n16,
                             unsigned *sink, hipStream_t st) {
  switch (unroll) {
    case 1: launch_probe_u<MODE, 1>(nt, blocks, in, out, n16, sink, st); break;
    case 2: launch_probe_u<MODE, 2>(nt, blocks, in, out, n16, sink, st); break;
    case 4: launch_probe_u<MODE, 4>(nt, blocks, in, out, n16, sink, st); break;
    case 8: launch_probe_u<MODE, 8>(nt, blocks, in, out, n16, sink, st); break;
    default: return -2;
  }
  return 0;
}

int launch_hbm_probe(int mode, const void *in, void *out, long long bytes, int blocks_per_cu, int unroll,
                     int nontemporal, unsigned *sink, hipStream_t st) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const long long n16 = bytes / 16;
  if (n16 <= 0) return 0;
  long long blocks = (long long)blocks_per_cu * cus;
  if (blocks > (n16 + 255) / 256) blocks = (n16 + 255) / 256;
  const uint4 *i4 = reinterpret_cast<const uint4 *>(in);
  uint4 *o4 = reinterpret_cast<uint4 *>(out);
  int rc = -2;
  switch (mode) {
    case 0: rc = launch_probe_mode<0>(unroll, nontemporal, (unsigned)blocks, i4, o4, n16, sink, st); break;
    case 1: rc = launch_probe_mode<1>(unroll, nontemporal, (unsigned)blocks, i4, o4, n16, sink, st); break;
    case 2: rc = launch_probe_mode<2>(unroll, nontemporal, (unsigned)blocks, i4, o4, n16, sink, st); break;
  }
  if (rc != 0) return rc;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

static int launch_copy_bytes(const uint8_t *in, uint8_t *out, long long bytes, hipStream_t st) {
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const long long n16 = bytes / 16;
  if (n16 > 0) {
    long long blocks = (n16 + 255) / 256;
    if (blocks > 8LL * cus) blocks = 8LL * cus;
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)blocks), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(in), reinterpret_cast<uint4 *>(out), n16);
  }
  const long long tail = bytes - n16 * 16;
  if (tail > 0) {
    CopyGeo g{};
    g.ndim = 1;
    g.esz = 1;
    g.shape[0] = tail;
    g.in_stride[0] = 1;
    g.in_offset = n16 * 16;
    g.total = tail;
    g.out_block = tail;
    g.out_pitch = tail;
    hipLaunchKernelGGL(k_strided_copy, dim3(1), dim3(256), 0, st, in, out + n16 * 16, g);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_strided_copy(const void *in, void *out, int elem_size, int ndim, const long long *shape,
                        const long long *in_strides, long long in_offset, long long out_block,
                        long long out_pitch, hipStream_t st) {
  if (ndim < 1 || ndim > 8) return -1;
  {
    // a C-contiguous view written contiguously is a plain byte copy
    bool contiguous = out_block <= 0;
    long long expect = elem_size, total = 1;
    for (int k = ndim - 1; k >= 0; --k) {
      if (shape[k] != 1 && in_strides[k] != expect) contiguous = false;
      expect *= shape[k];
      total *= shape[k];
    }
    const uint8_t *src = reinterpret_cast<const uint8_t *>(in) + in_offset;
    if (contiguous && ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(out)) % 16) == 0)
      return total == 0 ? 0 : launch_copy_bytes(src, reinterpret_cast<uint8_t *>(out), total * elem_size, st);
  }
  CopyGeo g;
  g.ndim = ndim;
  g.esz = elem_size;
  g.in_offset = in_offset;
  g.total = 1;
  for (int k = 0; k < ndim; ++k) {
    g.shape[k] = shape[k];
    g.in_stride[k] = in_strides[k];
    g.total *= shape[k];
  }
  if (g.total == 0) return 0;
  g.out_block = out_block > 0 ? out_block : g.total;
  g.out_pitch = out_block > 0 ? out_pitch : g.total;
  hipLaunchKernelGGL(k_strided_copy, dim3((unsigned)((g.total + 255) / 256)), dim3(256), 0, st,
                     reinterpret_cast<const uint8_t *>(in), reinterpret_cast<uint8_t *>(out), g);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// Broadcasting element-wise product (holonp::Multiply, multiply.cu:69-92): out has the
// broadcast shape; an operand's element stride is 0 on its broadcast axes.  Real operands
// are promoted to x + 0i when the output is complex.
struct MulGeo {
  int ndim;
  long long shape[8];
  long long sa[8], sb[8];  // element strides
  long long total;
  int a_complex, b_complex, out_complex;
};

__global__ void k_multiply(const void *a, const void *b, void *out, MulGeo g) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= g.total) return;
  long long rem = t, ia = 0, ib = 0;
  for (int k = g.ndim - 1; k >= 0; --k) {
    const long long idx = rem % g.shape[k];
    rem /= g.shape[k];
    ia += idx * g.sa[k];
    ib += idx * g.sb[k];
  }
  const float2 va = g.a_complex ? reinterpret_cast<const float2 *>(a)[ia]
                                : make_float2(reinterpret_cast<const float *>(a)[ia], 0.f);
  const float2 vb = g.b_complex ? reinterpret_cast<const float2 *>(b)[ib]
                                : make_float2(reinterpret_cast<const float *>(b)[ib], 0.f);
  if (g.out_complex)
    reinterpret_cast<float2 *>(out)[t] = cmul(va, vb);
  else
    reinterpret_cast<float *>(out)[t] = va.x * vb.x;
}

int launch_multiply(const void *a, int a_complex, const long long *sa, const void *b, int b_complex,
                    const long long *sb, int ndim, const long long *shape, void *out, hipStream_t st) {
  if (ndim < 1 || ndim > 8) return -1;
  MulGeo g;
  g.ndim = ndim;
  g.total = 1;
  for (int k = 0; k < ndim; ++k) {
    g.shape[k] = shape[k];
    g.sa[k] = sa[k];
    g.sb[k] = sb[k];
    g.total *= shape[k];
  }
  g.a_complex = a_complex;
  g.b_complex = b_complex;
  g.out_complex = a_complex || b_complex;
  if (g.total == 0) return 0;
  hipLaunchKernelGGL(k_multiply, dim3((unsigned)((g.total + 255) / 256)), dim3(256), 0, st, a, b, out, g);
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
  float scale;
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
  a.out[t] = make_float2(sr * a.scale, si * a.scale);
}

// Band DFT along one axis, tiled: a block owns 256 consecutive inner indices of one outer
// index, so every load of x[o, j, i0 .. i0+255] is one coalesced row; each thread keeps
// DFT_KC bins' accumulators and walks j once per chunk of bins.  The twiddles
// w[m] = exp(-2 pi i m / n) are built in LDS (FP64 sincospi, rounded once); the phase index
// (k j) mod n advances by k per sample (bins past the band advance by 0 and are not stored).
// DT: 1 U16, 2 F32, 3 CF32 (compile-time, so real inputs carry no imaginary terms).
constexpr int DFT_KC = 16, DFT_TILE = 256, DFT_MAXN = 4096;

template <int DT>
__global__ __launch_bounds__(DFT_TILE) void k_dft_axis_tiled(DftAxisArgs a, long long tiles) {
  extern __shared__ float2 tw[];
  for (int m = threadIdx.x; m < a.n; m += DFT_TILE) {
    double s, c;
    sincospi(2.0 * (double)m / (double)a.n, &s, &c);
    tw[m] = make_float2((float)c, (float)(-s));
  }
  __syncthreads();
  const long long o = blockIdx.x / tiles;
  const long long i = (blockIdx.x % tiles) * DFT_TILE + threadIdx.x;
  if (i >= a.inner) return;
  const long long nk = a.ke - a.kb;
  const long long base = o * (long long)a.n * a.inner + i;
  for (int k0 = a.kb; k0 < a.ke; k0 += DFT_KC) {
    float2 acc[DFT_KC];
    int m[DFT_KC], step[DFT_KC];
#pragma unroll
    for (int c = 0; c < DFT_KC; ++c) {
      acc[c] = make_float2(0.f, 0.f);
      m[c] = 0;
      step[c] = k0 + c < a.ke ? k0 + c : 0;
    }
#pragma unroll 2
    for (int j = 0; j < a.n; ++j) {
      const long long src = base + (long long)j * a.inner;
      float xr, xi = 0.f;
      if constexpr (DT == 1) {
        xr = (float)reinterpret_cast<const uint16_t *>(a.in)[src];
      } else if constexpr (DT == 2) {
        xr = reinterpret_cast<const float *>(a.in)[src];
      } else {
        const float2 v = reinterpret_cast<const float2 *>(a.in)[src];
        xr = v.x;
        xi = v.y;
      }
      float2 w[DFT_KC];
#pragma unroll
      for (int c = 0; c < DFT_KC; ++c) w[c] = tw[m[c]];
#pragma unroll
      for (int c = 0; c < DFT_KC; ++c) {
        if constexpr (DT == 3) {
          acc[c].x += xr * w[c].x - xi * w[c].y;
          acc[c].y += xr * w[c].y + xi * w[c].x;
        } else {
          acc[c].x += xr * w[c].x;
          acc[c].y += xr * w[c].y;
        }
        m[c] += step[c];
        m[c] = m[c] >= a.n ? m[c] - a.n : m[c];
      }
    }
#pragma unroll
    for (int c = 0; c < DFT_KC; ++c) {
      if (k0 + c < a.ke) {
        const long long k = k0 + c - a.kb;
        a.out[(o * nk + k) * a.inner + i] = make_float2(acc[c].x * a.scale, acc[c].y * a.scale);
      }
    }
  }
}

int launch_dft_axis(const void *in, int dtype, long long outer, int n, long long inner, int kb,
                    int ke, float scale, float2 *out, hipStream_t st) {
  DftAxisArgs a{in, dtype, outer, inner, n, kb, ke, scale, out};
  long long total = outer * (long long)(ke - kb) * inner;
  if (total == 0) return 0;
  const long long tiles = (inner + DFT_TILE - 1) / DFT_TILE;
  if (n <= DFT_MAXN && outer * tiles < (1LL << 31) && kb >= 0 && ke <= n) {
    // the phase counters reach at most (k0 + c) + n - 1 < 2 n before their wrap
    const dim3 grid((unsigned)(outer * tiles)), block(DFT_TILE);
    const size_t lds = (size_t)n * sizeof(float2);
    if (dtype == 1)
      hipLaunchKernelGGL(k_dft_axis_tiled<1>, grid, block, lds, st, a, tiles);
    else if (dtype == 2)
      hipLaunchKernelGGL(k_dft_axis_tiled<2>, grid, block, lds, st, a, tiles);
    else
      hipLaunchKernelGGL(k_dft_axis_tiled<3>, grid, block, lds, st, a, tiles);
  } else {
    hipLaunchKernelGGL(k_dft_axis, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, a);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hfx
