// Wave-level FFT engine: one 64-lane wavefront owns a whole transform of length
// N = 64 * E (E in {4, 8, 12, 16, 32}), E elements per lane.
//
// Convention: on entry lane l holds x[l + 64*m] in v[m]; on return it holds X[l + 64*m].
// Consecutive lanes therefore read and write consecutive indices (coalesced global
// access), and the data never leaves the wave:
//
//   pass 1 : radix-E DFT entirely in registers (butterfly j = lane, Stockham NS = 1)
//   pass k : radix-R Stockham pass through a wave-private LDS scratch (R <= 16), the
//            last pass leaving its results in registers in the natural layout.
//
// For N = 1024 this is radix 16 (registers) -> radix 16 -> radix 4: two LDS round trips
// and no workgroup barrier, against five barrier-separated radix-4 LDS passes in a
// workgroup-wide Stockham.  Within a wave, LDS instructions complete in issue order, so
// a write followed by a read of another lane's element needs only a compiler barrier.
//
// The scratch is padded by one element per 16 (pad(i) = i + i/16; one per 32 at E = 32), which
// makes the stride-E writes of pass 1 and the stride-NS accesses of the later passes free of
// bank conflicts (at E = 32 the stride of 34 elements that one per 16 gives would put lanes l
// and l + 16 on the same banks in pass 1; 33 spreads 32 lanes of 8-byte writes over all 64).  Pass twiddles come from an FP64-generated table tw[m] = exp(-2 pi i m/N);
// the small in-register DFTs use the exact constants cos/sin(2 pi m/32) and (2 pi m/12).
#pragma once

#include "hfx_device.hh"

namespace hfx {

// cos(2*pi*m/32), m = 0..8
__device__ constexpr float kCos32[9] = {1.0f,
                                        0.98078528040323044913f,
                                        0.92387953251128675613f,
                                        0.83146961230254523708f,
                                        0.70710678118654752440f,
                                        0.55557023301960222474f,
                                        0.38268343236508977173f,
                                        0.19509032201612826785f,
                                        0.0f};

__device__ __forceinline__ constexpr float cos32(int m) {  // cos(2*pi*m/32), any m >= 0
  m &= 31;
  if (m <= 8) return kCos32[m];
  if (m <= 16) return -kCos32[16 - m];
  if (m <= 24) return -kCos32[m - 16];
  return kCos32[32 - m];
}
__device__ __forceinline__ constexpr float sin32(int m) { return cos32(m + 24); }  // sin x = cos(x - pi/2)

__device__ __forceinline__ constexpr float cos12(int m) {  // cos(2*pi*m/12)
  constexpr float c[12] = {1.0f, 0.86602540378443864676f, 0.5f, 0.0f, -0.5f, -0.86602540378443864676f,
                           -1.0f, -0.86602540378443864676f, -0.5f, 0.0f, 0.5f, 0.86602540378443864676f};
  return c[m % 12];
}
__device__ __forceinline__ constexpr float sin12(int m) { return cos12(m + 9); }

// exp(sgn * 2*pi*i * p / R) for a compile-time (after unrolling) exponent p
template <int R> __device__ __forceinline__ float2 wr(int p, float sgn) {
  p %= R;
  if constexpr (32 % R == 0) {
    const int m = p * (32 / R);
    return make_float2(cos32(m), sgn * sin32(m));
  } else {
    static_assert(12 % R == 0, "unsupported in-register radix");
    const int m = p * (12 / R);
    return make_float2(cos12(m), sgn * sin12(m));
  }
}

// t * exp(S*2*pi*i*p/R) for a compile-time sign S (+1 inverse, -1 forward) and a
// compile-time exponent p after unrolling: quarter turns are exchanges and negations, the
// rest a complex multiply by the exact constant.
template <int R, int S> __device__ __forceinline__ float2 twr(float2 t, int p) {
  p %= R;
  if (p == 0) return t;
  if (4 * p == R) return S > 0 ? make_float2(-t.y, t.x) : make_float2(t.y, -t.x);      // * S i
  if (2 * p == R) return make_float2(-t.x, -t.y);                                       // * -1
  if (4 * p == 3 * R) return S > 0 ? make_float2(t.y, -t.x) : make_float2(-t.y, t.x);  // * -S i
  return cmul(t, wr<R>(p, (float)S));
}

// Radix-2/3/4 butterflies with the kernel sign fixed at compile time.
template <int R, int S> struct ButterflyS;
template <int S> struct ButterflyS<2, S> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float2 a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <int S> struct ButterflyS<4, S> {
  __device__ __forceinline__ static void run(float2 *v) {
    const float2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const float2 s13 = cadd(v[1], v[3]), d13 = twr<4, S>(csub(v[1], v[3]), 1);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};
template <int S> struct ButterflyS<3, S> {
  __device__ __forceinline__ static void run(float2 *v) {
    // w = exp(S*2*pi*i/3) = -1/2 + S*i*sqrt(3)/2
    const float c = -0.5f;
    const float s = (float)S * 0.86602540378443864676f;
    const float2 a = v[0], b = v[1], d = v[2];
    const float2 t = cadd(b, d);
    const float2 u = csub(b, d);
    const float2 m = make_float2(a.x + c * t.x, a.y + c * t.y);
    const float2 q = make_float2(-s * u.y, s * u.x);  // s * i * u
    v[0] = cadd(a, t);
    v[1] = cadd(m, q);
    v[2] = csub(m, q);
  }
};

// In-register DFT of length R (natural order in and out), kernel exp(S*2*pi*i*k*n/R).
// Composite lengths split as R = A*B (A = 4, 2 or 3):
//   Y[n2][k1] = W_R^{n2 k1} * DFT_A(x[B n1 + n2])[k1],  X[k1 + A k2] = DFT_B(Y[.][k1])[k2].
template <int R, int S> struct RegDft {
  __device__ __forceinline__ static void run(float2 (&v)[R]) {
    if constexpr (R == 2 || R == 3 || R == 4) {
      ButterflyS<R, S>::run(v);
    } else {
      constexpr int A = (R % 4 == 0) ? 4 : (R % 2 == 0) ? 2 : 3;
      constexpr int B = R / A;
      float2 y[B][A];
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) {
        float2 t[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
        ButterflyS<A, S>::run(t);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) y[n2][k1] = twr<R, S>(t[k1], n2 * k1);
      }
#pragma unroll
      for (int k1 = 0; k1 < A; ++k1) {
        float2 u[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) u[n2] = y[n2][k1];
        RegDft<B, S>::run(u);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
      }
    }
  }
};

__host__ __device__ constexpr int wave_pass_radix(int rem, int E) {
  // largest power of two <= 16 dividing both the remaining length and E (so that every
  // lane owns a whole number of butterflies)
  int r = 16;
  while (r > 1 && (rem % r != 0 || E % r != 0)) r /= 2;
  return r;
}

template <int N> struct WaveFft {
  static_assert(N % 64 == 0, "wave FFT length must be a multiple of 64");
  static constexpr int E = N / 64;
  static constexpr int PSH = E >= 32 ? 5 : 4;  // one pad element per 2^PSH
  static constexpr int SCR = N + (N >> PSH);    // padded scratch elements per transform
  // Pass twiddles are stored pass-major: the pass with stride NS and radix R owns
  // (R-1)*NS entries W_{NS*R}^{r*k} at [r-1][k], so the lanes of a wave (consecutive k)
  // read consecutive entries.  Over all passes this is N - E entries.
  static constexpr int TW = N - E;

  __device__ __forceinline__ static int pad(int i) { return i + (i >> PSH); }

  // offset of the table of the pass that starts at stride NS
  static constexpr int tw_offset(int ns) {
    int off = 0;
    for (int s = E; s < ns;) {
      const int r = wave_pass_radix(N / s, E);
      off += (r - 1) * s;
      s *= r;
    }
    return off;
  }

  // Builds the pass-major table from an N-entry table t[m] = exp(-2 pi i m / N).
  __device__ static void build_twiddles(float2 *dst, const float2 *t, int tid, int nthreads) {
    for (int s = E; s < N;) {
      const int r = wave_pass_radix(N / s, E);
      const int off = tw_offset(s);
      for (int i = tid; i < (r - 1) * s; i += nthreads) {
        const int rr = i / s + 1, k = i % s;
        dst[off + i] = t[rr * k * (N / (s * r))];
      }
      s *= r;
    }
  }

  template <int NS, int S>
  __device__ __forceinline__ static void pass(float2 (&v)[E], float2 *scr, const float2 *tw, int lane) {
    constexpr int R = wave_pass_radix(N / NS, E);
    static_assert(R >= 2, "bad factorisation");
    constexpr int NB = N / R;
    constexpr int PER = NB / 64;  // butterflies per lane
    constexpr bool LAST = NS * R == N;
    constexpr int TOFF = tw_offset(NS);
    float2 u[PER][R];
    float2 w[PER][R];  // w[q][0] unused
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 0; r < R; ++r) u[q][r] = scr[pad(lane + 64 * q + r * NB)];
    // all twiddle reads issued together with the data reads, so their latencies overlap
    // instead of one LDS round trip per pair; the empty asm keeps the compiler from sinking
    // each read down to its use
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 1; r < R; ++r) w[q][r] = tw[TOFF + (r - 1) * NS + (lane + 64 * q) % NS];
#pragma unroll
    for (int q = 0; q < PER; ++q)
#pragma unroll
      for (int r = 1; r < R; ++r) asm volatile("" : "+v"(w[q][r].x), "+v"(w[q][r].y));
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
#pragma unroll
      for (int r = 1; r < R; ++r)
        // the table holds the forward twiddles exp(-2 pi i m/N); the inverse conjugates them
        u[q][r] = S < 0 ? cmul(u[q][r], w[q][r]) : cmulc(u[q][r], w[q][r]);
      RegDft<R, S>::run(u[q]);
    }
    if constexpr (LAST) {
      // outputs j + r*NS with NS = N/R: lane holds X[lane + 64*(q + r*E/R)]
#pragma unroll
      for (int q = 0; q < PER; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) v[q + r * (E / R)] = u[q][r];
    } else {
#pragma unroll
      for (int q = 0; q < PER; ++q) {
        const int j = lane + 64 * q;
        const int base = (j / NS) * NS * R + j % NS;
#pragma unroll
        for (int r = 0; r < R; ++r) scr[pad(base + r * NS)] = u[q][r];
      }
      __builtin_amdgcn_wave_barrier();
      pass<NS * R, S>(v, scr, tw, lane);
    }
  }

  // Transforms v in place (see the convention above); scr: SCR wave-private LDS elements;
  // tw: the TW-entry pass-major table from build_twiddles; inv: kernel exp(+...), unscaled.
  template <bool INV>
  __device__ __forceinline__ static void run(float2 (&v)[E], float2 *scr, const float2 *tw, int lane) {
    stage_in<INV>(v, scr, lane);
    stage_out<INV>(v, scr, tw, lane);
  }
  // The two halves of run(), for callers that load the next transform's input while this one
  // runs: stage_in consumes v (pass 1, into the scratch), after which v may be reloaded;
  // stage_out runs the remaining passes and leaves X in o.
  template <bool INV> __device__ __forceinline__ static void stage_in(float2 (&v)[E], float2 *scr, int lane) {
    constexpr int S = INV ? 1 : -1;
    RegDft<E, S>::run(v);  // pass 1: butterfly j = lane over x[lane + 64 r]
#pragma unroll
    for (int r = 0; r < E; ++r) scr[pad(lane * E + r)] = v[r];
    __builtin_amdgcn_wave_barrier();
  }
  template <bool INV>
  __device__ __forceinline__ static void stage_out(float2 (&o)[E], float2 *scr, const float2 *tw, int lane) {
    pass<E, INV ? 1 : -1>(o, scr, tw, lane);
  }
  // run-time direction (a wave-uniform branch between the two instantiations)
  __device__ __forceinline__ static void run(float2 (&v)[E], float2 *scr, const float2 *tw, int lane,
                                             bool inv) {
    if (inv)
      run<true>(v, scr, tw, lane);
    else
      run<false>(v, scr, tw, lane);
  }
};

}  // namespace hfx
