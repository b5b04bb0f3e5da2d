// Band-limited time DFT of a uint16 frame window, run on the MFMA units of gfx950.
//
// For every pixel p the path needs S_k(p) = sum_n x[n](p) * exp(-2*pi*i*k*n/T) for the band
// bins k in [kb, kb+nbc) only (rfft.cu:148-171 computes all T/2+1 bins and slices them
// afterwards, graph_builder.cc:259-266).  Written as a contraction it is
//     D[2*nbc x P] = Wm[2*nbc x T] * X[T x P],   Wm rows = (cos, -sin) per bin,
// i.e. a skinny GEMM with K = T.  One wave owns 64 pixels, split into four 16-pixel
// sub-tiles {base + 4*j + t : j < 16} so that each lane loads 4 adjacent pixels of one
// frame with a single 8-byte load; the four values feed the B operand of four MFMAs.
//
// The samples are centred, x - 2048, before the products.  For 12-bit data this is exact
// in FP32 and FP16 (|x - 2048| <= 2048 fits the 11-bit FP16 significand), and it changes
// no bin k != 0 (sum_n exp(-2*pi*i*k*n/T) = 0); bin 0 gets 2048*T added back.
//
// Precisions (accumulation is FP32 in all of them):
//   FP32       : v_mfma_f32_16x16x4_f32, operands FP32.
//   FP16       : v_mfma_f32_16x16x32_f16, samples exact, twiddles rounded to FP16.
//   BF16       : v_mfma_f32_16x16x32_bf16, samples and twiddles rounded to BF16.
//   FP16_SPLIT : two FP16 MFMAs per tile, twiddle = hi + lo (both FP16), samples exact.
//
// Epilogue: optional Fresnel lens multiply (separable chirp lx[x]*ly[y]) and, for the
// two-sided FFT path, the conjugate planes of the negative band.
#include "hfx_kernels.hh"

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

namespace hfx {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

struct DftArgs {
  const uint16_t *ring;  // [ring_frames][npix]
  long long ring_frames;
  long long first;       // first frame of the window, in [0, ring_frames)
  int T;
  long long npix;
  int W;
  int kb, nb, nbc;       // band start, planes per sideband, computed bins (nb or nb+1)
  int both;              // write negative-band conjugate planes at [nb, 2nb)
  float2 *planes;        // [planes][npix]
  const float2 *lx, *ly; // Fresnel lens vectors or null
};

__device__ __forceinline__ void dft_store(const DftArgs &a, int bin, long long p0, const float *re,
                                          const float *im) {
  // bin: computed bin index (kb + bin is the frequency index); p0: first of 4 pixels.
  float2 v[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) v[t] = make_float2(re[t], im[t]);
  if (a.kb + bin == 0) {
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t].x += 2048.0f * (float)a.T;
  }
  if (p0 >= a.npix) return;
  float2 L[4];
  bool lens = a.lx != nullptr;
  if (lens) {
    int y = (int)(p0 / a.W), x = (int)(p0 % a.W);  // W % 4 == 0: 4 pixels share a row
    float2 ly = a.ly[y];
#pragma unroll
    for (int t = 0; t < 4; ++t) L[t] = cmul(a.lx[x + t], ly);
  }
  if (bin < a.nb) {
    float2 o[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) o[t] = lens ? cmul(v[t], L[t]) : v[t];
    float4 *dst = reinterpret_cast<float4 *>(a.planes + (long long)bin * a.npix + p0);
    dst[0] = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
    dst[1] = make_float4(o[2].x, o[2].y, o[3].x, o[3].y);
  }
  if (a.both) {
    int j = a.nb - bin;  // negative-band plane holding conj(X[kb+bin]) = X[T-(kb+bin)]
    if (j >= 0 && j < a.nb) {
      float2 o[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        float2 c = make_float2(v[t].x, -v[t].y);
        o[t] = lens ? cmul(c, L[t]) : c;
      }
      float4 *dst = reinterpret_cast<float4 *>(a.planes + (long long)(a.nb + j) * a.npix + p0);
      dst[0] = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
      dst[1] = make_float4(o[2].x, o[2].y, o[3].x, o[3].y);
    }
  }
}

__device__ __forceinline__ uint2 load4(const DftArgs &a, long long frame, long long p0) {
  if (p0 >= a.npix) return make_uint2(0x08000800u, 0x08000800u);  // 2048: centred zero
  const uint16_t *src = a.ring + frame * a.npix + p0;
  return *reinterpret_cast<const uint2 *>(src);
}

// ---------------------------------------------------------------- FP32 MFMA 16x16x4
template <int NBT>
__global__ __launch_bounds__(256) void k_dft_f32(DftArgs a, const float *wm) {
  extern __shared__ float lds_w[];  // [NBT*16][T+1]
  const int T = a.T, TP = T + 1;
  for (int i = threadIdx.x; i < NBT * 16 * T; i += blockDim.x) {
    int r = i / T, n = i % T;
    lds_w[r * TP + n] = wm[i];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long base = ((long long)blockIdx.x * 4 + wave) * 64;
  const long long p0 = base + 4 * (lane & 15);
  const int sub = lane >> 4;  // k index inside the MFMA step / row group of the output

  f4 acc[NBT][4];
#pragma unroll
  for (int bt = 0; bt < NBT; ++bt)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[bt][t] = f4{0.f, 0.f, 0.f, 0.f};

  long long frame = a.first + sub;
  if (frame >= a.ring_frames) frame -= a.ring_frames;
  for (int s = 0; s < T / 4; ++s) {
    uint2 raw = load4(a, frame, p0);
    float x[4];
    x[0] = (float)(raw.x & 0xffffu) - 2048.0f;
    x[1] = (float)(raw.x >> 16) - 2048.0f;
    x[2] = (float)(raw.y & 0xffffu) - 2048.0f;
    x[3] = (float)(raw.y >> 16) - 2048.0f;
    const int n = 4 * s + sub;
#pragma unroll
    for (int bt = 0; bt < NBT; ++bt) {
      float w = lds_w[(bt * 16 + (lane & 15)) * TP + n];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        acc[bt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w, x[t], acc[bt][t], 0, 0, 0);
    }
    frame += 4;
    if (frame >= a.ring_frames) frame -= a.ring_frames;
  }

  // D[row][pixel]: row = bt*16 + 4*sub + r, pixel = base + 4*(lane&15) + t
#pragma unroll
  for (int bt = 0; bt < NBT; ++bt) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int bin = bt * 8 + 2 * sub + h;
      if (bin < a.nbc) {
        float re[4], im[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          re[t] = acc[bt][t][2 * h];
          im[t] = acc[bt][t][2 * h + 1];
        }
        dft_store(a, bin, p0, re, im);
      }
    }
  }
}

// ---------------------------------------------------------------- 16x16x32 FP16 / BF16
// MODE 0: FP16, 1: BF16, 2: FP16 hi/lo split twiddles.
template <int NBT, int MODE>
__global__ __launch_bounds__(256) void k_dft_lp(DftArgs a, const uint16_t *wm_hi,
                                                const uint16_t *wm_lo) {
  // operands stored as raw 16-bit words; row stride padded by 8 elements
  extern __shared__ uint16_t lds16[];
  const int T = a.T, TP = T + 8;
  uint16_t *lhi = lds16;
  uint16_t *llo = lds16 + NBT * 16 * TP;
  for (int i = threadIdx.x; i < NBT * 16 * T; i += blockDim.x) {
    int r = i / T, n = i % T;
    lhi[r * TP + n] = wm_hi[i];
    if (MODE == 2) llo[r * TP + n] = wm_lo[i];
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const long long base = ((long long)blockIdx.x * 4 + wave) * 64;
  const long long p0 = base + 4 * (lane & 15);
  const int sub = lane >> 4;

  f4 acc[NBT][4];
#pragma unroll
  for (int bt = 0; bt < NBT; ++bt)
#pragma unroll
    for (int t = 0; t < 4; ++t) acc[bt][t] = f4{0.f, 0.f, 0.f, 0.f};

  for (int s = 0; s < T / 32; ++s) {
    // lane holds frames n = 32*s + 8*sub + q, q < 8, for its 4 pixels
    uint2 raw[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      long long frame = a.first + 32 * s + 8 * sub + q;
      if (frame >= a.ring_frames) frame -= a.ring_frames;
      raw[q] = load4(a, frame, p0);
    }
    h8 xh[4];
    bf8 xb[4];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      float v0 = (float)(raw[q].x & 0xffffu) - 2048.0f;
      float v1 = (float)(raw[q].x >> 16) - 2048.0f;
      float v2 = (float)(raw[q].y & 0xffffu) - 2048.0f;
      float v3 = (float)(raw[q].y >> 16) - 2048.0f;
      if (MODE == 1) {
        xb[0][q] = (__bf16)v0;
        xb[1][q] = (__bf16)v1;
        xb[2][q] = (__bf16)v2;
        xb[3][q] = (__bf16)v3;
      } else {
        xh[0][q] = (_Float16)v0;
        xh[1][q] = (_Float16)v1;
        xh[2][q] = (_Float16)v2;
        xh[3][q] = (_Float16)v3;
      }
    }
    const int k0 = 32 * s + 8 * sub;
#pragma unroll
    for (int bt = 0; bt < NBT; ++bt) {
      const int row = bt * 16 + (lane & 15);
      const uint4 whi = *reinterpret_cast<const uint4 *>(&lhi[row * TP + k0]);
      if (MODE == 1) {
        bf8 w = __builtin_bit_cast(bf8, whi);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[t], acc[bt][t], 0, 0, 0);
      } else {
        h8 w = __builtin_bit_cast(h8, whi);
#pragma unroll
        for (int t = 0; t < 4; ++t)
          acc[bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w, xh[t], acc[bt][t], 0, 0, 0);
        if (MODE == 2) {
          const uint4 wlo = *reinterpret_cast<const uint4 *>(&llo[row * TP + k0]);
          h8 wl = __builtin_bit_cast(h8, wlo);
#pragma unroll
          for (int t = 0; t < 4; ++t)
            acc[bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wl, xh[t], acc[bt][t], 0, 0, 0);
        }
      }
    }
  }

#pragma unroll
  for (int bt = 0; bt < NBT; ++bt) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int bin = bt * 8 + 2 * sub + h;
      if (bin < a.nbc) {
        float re[4], im[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          re[t] = acc[bt][t][2 * h];
          im[t] = acc[bt][t][2 * h + 1];
        }
        dft_store(a, bin, p0, re, im);
      }
    }
  }
}

template <int NBT>
static int launch_nbt(const DftLaunch &L, hipStream_t st) {
  DftArgs a;
  a.ring = L.ring;
  a.ring_frames = L.ring_frames;
  a.first = L.first;
  a.T = L.T;
  a.npix = L.npix;
  a.W = L.W;
  a.kb = L.kb;
  a.nb = L.nb;
  a.nbc = L.nbc;
  a.both = L.both;
  a.planes = L.planes;
  a.lx = L.lx;
  a.ly = L.ly;
  long long waves = (L.npix + 63) / 64;
  dim3 grid((unsigned)((waves + 3) / 4));
  if (L.precision == 0) {
    size_t lds = (size_t)NBT * 16 * (L.T + 1) * sizeof(float);
    hipLaunchKernelGGL(k_dft_f32<NBT>, grid, dim3(256), lds, st, a, L.wm_f32);
  } else {
    size_t lds = (size_t)NBT * 16 * (L.T + 8) * 2 * (L.precision == 3 ? 2 : 1);
    const uint16_t *hi = L.precision == 2 ? L.wm_bf16 : L.wm_f16hi;
    if (L.precision == 1)
      hipLaunchKernelGGL((k_dft_lp<NBT, 0>), grid, dim3(256), lds, st, a, hi, L.wm_f16lo);
    else if (L.precision == 2)
      hipLaunchKernelGGL((k_dft_lp<NBT, 1>), grid, dim3(256), lds, st, a, hi, L.wm_f16lo);
    else
      hipLaunchKernelGGL((k_dft_lp<NBT, 2>), grid, dim3(256), lds, st, a, hi, L.wm_f16lo);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_time_dft(const DftLaunch &L, hipStream_t st) {
  int nbt = (2 * L.nbc + 15) / 16;
  switch (nbt) {
    case 1: return launch_nbt<1>(L, st);
    case 2: return launch_nbt<2>(L, st);
    case 3: return launch_nbt<3>(L, st);
    case 4: return launch_nbt<4>(L, st);
    case 5: return launch_nbt<5>(L, st);
    case 6: return launch_nbt<6>(L, st);
    case 7: return launch_nbt<7>(L, st);
    case 8: return launch_nbt<8>(L, st);
    default: return -2;
  }
}

}  // namespace hfx
