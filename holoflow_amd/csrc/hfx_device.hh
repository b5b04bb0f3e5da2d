// Device-side building blocks shared by the gfx950 kernels of the Doppler-moment path:
// complex helpers, small radix butterflies and an LDS-staged Stockham FFT engine.
//
// The Stockham formulation used here is the self-sorting variant of Cooley-Tukey
// (T. G. Stockham 1966; the GPU form of Govindaraju et al., SC'08): at pass p with
// NS = product of the radices already applied, butterfly j reads R elements at
// j + r*N/R, twiddles element r by exp(-+2*pi*i*r*(j%NS)/(NS*R)), runs an R-point DFT and
// writes them to (j/NS)*NS*R + j%NS + r*NS.  After the last pass the data is in natural
// order, so no bit reversal is needed.  Each pass keeps a thread's butterflies in
// registers between two workgroup barriers, so one LDS buffer is enough.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hfx {

struct __align__(8) c32 {
  float x, y;
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * b with its rounding pinned (one product and one fused multiply-add per part): for
// products that must round alike in every inlined copy, whatever the compiler contracts
__device__ __forceinline__ float2 cmul_rn(float2 a, float2 b) {
  return make_float2(__fmaf_rn(a.x, b.x, -__fmul_rn(a.y, b.y)), __fmaf_rn(a.x, b.y, __fmul_rn(a.y, b.x)));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
  return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// sgn * i * a: sgn = -1 for the forward kernel (multiply by -i), +1 for the inverse.
__device__ __forceinline__ float2 mul_si(float2 a, float sgn) {
  return make_float2(-sgn * a.y, sgn * a.x);
}

// R-point DFT in registers with kernel exp(sgn*2*pi*i*k*n/R), sgn = -1 forward, +1 inverse.
template <int R> struct Butterfly;

template <> struct Butterfly<2> {
  __device__ __forceinline__ static void run(float2 *v, float) {
    float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <> struct Butterfly<4> {
  __device__ __forceinline__ static void run(float2 *v, float sgn) {
    float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    float2 s13 = cadd(v[1], v[3]), d13 = mul_si(csub(v[1], v[3]), sgn);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <> struct Butterfly<3> {
  __device__ __forceinline__ static void run(float2 *v, float sgn) {
    // w = exp(sgn*2*pi*i/3) = -1/2 + sgn*i*sqrt(3)/2
    const float c = -0.5f;
    const float s = sgn * 0.86602540378443864676f;
    float2 a = v[0], b = v[1], d = v[2];
    float2 t = cadd(b, d);
    float2 u = csub(b, d);
    float2 m = make_float2(a.x + c * t.x, a.y + c * t.y);
    float2 q = make_float2(-s * u.y, s * u.x);  // s * i * u
    v[0] = cadd(a, t);
    v[1] = cadd(m, q);
    v[2] = csub(m, q);
  }
};

// Radix chosen for the remaining length REM = N/NS: fours first, then a two, then threes.
__host__ __device__ constexpr int pick_radix(int rem) {
  return (rem % 4 == 0) ? 4 : (rem % 2 == 0) ? 2 : (rem % 3 == 0) ? 3 : 0;
}
__host__ __device__ constexpr bool fft_size_ok(int n) {
  if (n < 2) return false;
  while (n % 4 == 0) n /= 4;
  while (n % 2 == 0) n /= 2;
  while (n % 3 == 0) n /= 3;
  return n == 1;
}

// Layout of a batch of ROWS transforms of length N inside LDS:
//   element (row, i) lives at lds[row * RS + i * ES].
// Row-major batches use RS = N (+pad), ES = 1; column batches use RS = 1, ES = ROWS.
// Butterflies are enumerated so that consecutive lanes touch consecutive LDS words:
//   ES == 1 : b -> (row = b / NB, j = b % NB)
//   ES  > 1 : b -> (row = b % ROWS, j = b / ROWS)
// `tid` is the thread index in [0, NT); callers pass a value the compiler cannot treat as
// loop-invariant so that index arithmetic is not hoisted across a surrounding plane loop.
template <int N, int ROWS, int RS, int ES, int NT> struct LdsFft {
  static_assert(fft_size_ok(N), "FFT length must be 2^a * 3^b");

  template <int NS>
  __device__ __forceinline__ static void pass(float2 *lds, const float2 *tw /* N entries */,
                                              int tid, float sgn) {
    constexpr int R = pick_radix(N / NS);
    constexpr int NB = N / R;
    constexpr int TOT = ROWS * NB;
    constexpr int PER = (TOT + NT - 1) / NT;
    float2 v[PER][R];
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int b = tid + it * NT;
      if ((TOT % NT == 0) || b < TOT) {
        const int row = (ES == 1) ? b / NB : b % ROWS;
        const int j = (ES == 1) ? b % NB : b / ROWS;
#pragma unroll
        for (int r = 0; r < R; ++r) v[it][r] = lds[row * RS + (j + r * NB) * ES];
      }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < PER; ++it) {
      const int b = tid + it * NT;
      if ((TOT % NT == 0) || b < TOT) {
        const int row = (ES == 1) ? b / NB : b % ROWS;
        const int j = (ES == 1) ? b % NB : b / ROWS;
        const int k = j % NS;
        if (NS > 1) {
#pragma unroll
          for (int r = 1; r < R; ++r) {
            float2 w = tw[r * k * (N / (NS * R))];
            w.y = -sgn * w.y;  // table holds exp(-2 pi i m/N); the inverse uses its conjugate
            v[it][r] = cmul(v[it][r], make_float2(w.x, w.y));
          }
        }
        Butterfly<R>::run(v[it], sgn);
        const int base = (j / NS) * NS * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[row * RS + (base + r * NS) * ES] = v[it][r];
      }
    }
    __syncthreads();
    if constexpr (NS * R < N) pass<NS * R>(lds, tw, tid, sgn);
  }

  // Runs the full transform (inv: kernel exp(+2 pi i k n / N), unnormalised).  The caller
  // must have synchronised after filling `lds` and `tw`.
  __device__ __forceinline__ static void run(float2 *lds, const float2 *tw, int tid, bool inv) {
    pass<1>(lds, tw, tid, inv ? 1.0f : -1.0f);
  }
};

// Bijective workgroup -> tile map that gives each of the 8 XCDs (workgroups are dispatched
// round-robin over them) a contiguous run of tiles, so neighbouring rows written into the
// column-blocked layout share an L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
  const int n8 = n & ~7;
  if (b >= n8) return b;
  return (b & 7) * (n8 >> 3) + (b >> 3);
}

// |v|^2 for the moment sums with its rounding pinned (one product, one fused multiply-add), so
// that every pass and build forms the sums from the same values
__device__ __forceinline__ float sq_mag(float2 v) { return __fmaf_rn(v.x, v.x, __fmul_rn(v.y, v.y)); }

// |X| for the moment sums: the hardware square root (v_sqrt_f32, within 1 ulp) instead of the
// correctly rounded sqrtf sequence, which costs about ten more VALU instructions per element in
// the column pass; the moments are sums over planes held to a relative tolerance, not bits.
__device__ __forceinline__ float mag_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Column-blocked layout of an H x W plane with column blocks of C in {1, 2, 4}: element
// (y, x) sits at (x / C) * H * C + y * C + x % C.  For the wave-owned row x = lane + 64 m
// (C divides 64) that is blk_lane_offset(lane) + y * C + m * 64 * H, so a row store or load
// needs one shift-and-mask per lane instead of a division per element.
__device__ __forceinline__ long long blk_lane_offset(int lane, int H, int C) {
  const unsigned ul = (unsigned)lane, lc = (unsigned)__builtin_ctz((unsigned)C);
  return (long long)(ul >> lc) * H * C + (ul & (unsigned)(C - 1));
}

}  // namespace hfx
