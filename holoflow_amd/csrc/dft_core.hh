// MFMA building block of the band-limited time DFT (shared by the stand-alone time-DFT
// kernel and the fused time-DFT + row-FFT kernel).
//
// For 64 pixels owned by one wave it computes D[2*nbc x 64] = Wm[2*nbc x T] * X[T x 64],
// the rows of Wm being (cos, -sin)(2*pi*(kb+b)*n/T) of the band bins b (rfft.cu:148-171
// computes all T/2+1 bins and graph_builder.cc:259-266 slices the band afterwards).  The
// 64 pixels are split into four 16-pixel sub-tiles {p0 + 4*j + t : j < 16} so a lane loads
// 4 adjacent pixels of one frame with a single 8-byte load; the four values feed the B
// operand of four MFMAs.
//
// Samples are centred, x - 2048, before the products.  For 12-bit data this is exact in
// FP32 and FP16 (|x - 2048| <= 2048 fits the 11-bit FP16 significand) and changes no bin
// k != 0 (sum_n exp(-2*pi*i*k*n/T) = 0); bin 0 gets 2048*T added back by the caller.
//
// MODE 0 FP32      : v_mfma_f32_16x16x4_f32, FP32 operands
// MODE 1 FP16      : v_mfma_f32_16x16x32_f16, exact samples, twiddles rounded to FP16
// MODE 2 BF16      : v_mfma_f32_16x16x32_bf16, samples and twiddles rounded to BF16
// MODE 3 FP16 split: two FP16 MFMAs per tile with twiddle = hi + lo, exact samples
// Accumulation is FP32 in every mode.
//
// Output layout (MFMA D registers): acc[bt][t][r] holds row bt*16 + 4*(lane>>4) + r for
// pixel p0 + 4*(lane&15) + t; rows 2b / 2b+1 are Re / Im of bin b, so a lane holds bins
// bt*8 + 2*(lane>>4) + {0, 1}.
#pragma once

#include <type_traits>

#include "hfx_device.hh"

namespace hfx {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

struct DftRing {
  const uint16_t *ring;   // [ring_frames][npix]
  long long ring_frames;
  long long first;        // first frame of the window, in [0, ring_frames)
  long long npix;
  int T;
};

// Frame loads are non-temporal: a window's samples are read once per launch, and keeping them
// out of the caches leaves the last-level cache to the band planes that the column pass reads
// back right after (whole-batch throughput +3 % at 1024^2, T = 64; results unchanged).
// Loads carry no bounds check: a check is a branch around every load, which splits a step loop
// into basic blocks and makes a step wait for every load in flight instead of its own (front
// kernel 92 -> 86 us at 1024^2, T = 64).  Callers clamp their pixel indices instead
// (dft_px): pixels past the end of a ragged last tile read the frame's last pixels, and the
// caller stores no result for them (the MFMA output column of a pixel depends on that pixel's
// inputs alone).
// Frame and pixel counts are below 2^31 (the plan and ring checks enforce it): the element
// offset is one 32 x 32 -> 64-bit multiply-add, where 64-bit index arithmetic costs about
// three times the VALU instructions in the hottest loop of the time DFT.
__device__ __forceinline__ uint2 dft_load4(const DftRing &r, int frame, long long p) {
  typedef unsigned u2v __attribute__((ext_vector_type(2)));
  const unsigned long long off = (unsigned long long)(unsigned)frame * (unsigned)r.npix + (unsigned long long)p;
  const u2v v = __builtin_nontemporal_load(reinterpret_cast<const u2v *>(r.ring + off));
  return make_uint2(v.x, v.y);
}

// First of the 4 pixels a lane loads in tile i of a wave's tiles at p0w, clamped into the frame
__device__ __forceinline__ long long dft_px(const DftRing &r, long long p0w, int lane, int i) {
  const long long p = p0w + 64 * i + 4 * (lane & 15);
  return p < r.npix ? p : r.npix - 4;
}

// Stages the operand matrix into LDS.  FP32 rows (MODE 0) use a pitch of T+4 floats, so the
// 16 rows x 4 frame offsets read by one MFMA step fall in 64 distinct banks; 16-bit rows
// (MODE 1..3) use T+8 elements (16-byte aligned row starts; `lo` rows follow the `hi`
// rows for MODE 3).
template <int NBT, int MODE>
__device__ __forceinline__ void dft_stage_matrix(void *lds, const void *wm_hi, const void *wm_lo,
                                                 int T, int tid, int nthreads) {
  if (MODE == 0) {
    float *l = reinterpret_cast<float *>(lds);
    const float *w = reinterpret_cast<const float *>(wm_hi);
    for (int i = tid; i < NBT * 16 * T; i += nthreads) l[(i / T) * (T + 4) + i % T] = w[i];
  } else {
    uint16_t *l = reinterpret_cast<uint16_t *>(lds);
    const uint16_t *hi = reinterpret_cast<const uint16_t *>(wm_hi);
    const uint16_t *lo = reinterpret_cast<const uint16_t *>(wm_lo);
    const int TP = T + 8;
    for (int i = tid; i < NBT * 16 * T; i += nthreads) {
      l[(i / T) * TP + i % T] = hi[i];
      if (MODE == 3) l[NBT * 16 * TP + (i / T) * TP + i % T] = lo[i];
    }
  }
}

__host__ __device__ constexpr size_t dft_matrix_lds_bytes(int nbt, int mode, int T) {
  return mode == 0 ? (size_t)nbt * 16 * (T + 4) * 4 : (size_t)nbt * 16 * (T + 8) * 2 * (mode == 3 ? 2 : 1);
}

// Loads a wave issues ahead of its MFMA loop: the first 32 frames (MODE 0: 8 steps of 4
// frames; MODE 1..3: one step of 32), for each of its TPW tiles.
template <int MODE> constexpr int dft_pre_steps() { return 8; }

__device__ __forceinline__ int ring_frame(const DftRing &r, int n) {
  // unsigned: first + n may pass 2^31 - 1 for a ring of nearly 2^31 frames
  const unsigned f = (unsigned)r.first + (unsigned)n, R = (unsigned)r.ring_frames;
  return (int)(f >= R ? f - R : f);  // n < T <= ring_frames
}

// The wave owns TPW adjacent 64-pixel tiles starting at p0w (a multiple of 4); tile i
// covers p0w + 64*i + {4*j + t}.  npix % 4 == 0.
// PRE (FP32 only) lowers the number of prefetched steps where registers are scarce.
template <int MODE, int TPW, int PRE = dft_pre_steps<MODE>()>
__device__ __forceinline__ void dft_prefetch(const DftRing &r, long long p0w, int lane,
                                             uint2 (&pre)[TPW][PRE]) {
  static_assert(MODE == 0 || PRE == 8, "16-bit modes prefetch one 32-frame step");
  const int sub = lane >> 4;
  long long p0[TPW];  // clamped (dft_px)
#pragma unroll
  for (int i = 0; i < TPW; ++i) p0[i] = dft_px(r, p0w, lane, i);
  if (MODE == 0) {
#pragma unroll
    for (int s = 0; s < PRE; ++s) {
      if (4 * s < r.T) {
        const int f = ring_frame(r, 4 * s + sub);
#pragma unroll
        for (int i = 0; i < TPW; ++i) pre[i][s] = dft_load4(r, f, p0[i]);
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int f = ring_frame(r, 8 * sub + q);
#pragma unroll
      for (int i = 0; i < TPW; ++i) pre[i][q] = dft_load4(r, f, p0[i]);
    }
  }
}

__device__ __forceinline__ void unpack4(uint2 raw, float (&x)[4]) {
  x[0] = (float)(raw.x & 0xffffu) - 2048.0f;
  x[1] = (float)(raw.x >> 16) - 2048.0f;
  x[2] = (float)(raw.y & 0xffffu) - 2048.0f;
  x[3] = (float)(raw.y >> 16) - 2048.0f;
}

// MFMA loop over the window, starting from the prefetched first 32 frames `pre` and
// loading the rest one batch ahead of its use.
template <int NBT, int MODE, int TPW, int PRE = dft_pre_steps<MODE>(), int KB = 8 / TPW>
__device__ __forceinline__ void dft_tile(const DftRing &r, const void *lds, long long p0w, int lane,
                                         const uint2 (&pre)[TPW][PRE],
                                         f4 (&acc)[TPW][NBT][4]) {
  const int T = r.T;
  const int sub = lane >> 4;
  long long p0[TPW];  // clamped (dft_px)
#pragma unroll
  for (int i = 0; i < TPW; ++i) p0[i] = dft_px(r, p0w, lane, i);
#pragma unroll
  for (int i = 0; i < TPW; ++i)
#pragma unroll
    for (int bt = 0; bt < NBT; ++bt)
#pragma unroll
      for (int t = 0; t < 4; ++t) acc[i][bt][t] = f4{0.f, 0.f, 0.f, 0.f};

  if (MODE == 0) {
    const float *lw = reinterpret_cast<const float *>(lds);
    const int TP = T + 4;
    auto step = [&](const uint2 (&raw)[TPW], int s) {
      const int n = 4 * s + sub;
      float x[TPW][4];
#pragma unroll
      for (int i = 0; i < TPW; ++i) unpack4(raw[i], x[i]);
#pragma unroll
      for (int bt = 0; bt < NBT; ++bt) {
        const float w = lw[(bt * 16 + (lane & 15)) * TP + n];
#pragma unroll
        for (int i = 0; i < TPW; ++i)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            acc[i][bt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w, x[i][t], acc[i][bt][t], 0, 0, 0);
      }
    };
    const int steps = T / 4;
    // steps beyond the prefetched PRE come in batches of K, one batch ahead
    constexpr int K = KB;
    uint2 nxt[K][TPW];
    auto load_batch = [&](int s0) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        if (s0 + k < steps) {
          const int f = ring_frame(r, 4 * (s0 + k) + sub);
#pragma unroll
          for (int i = 0; i < TPW; ++i) nxt[k][i] = dft_load4(r, f, p0[i]);
        }
      }
    };
    if (steps > PRE) load_batch(PRE);
#pragma unroll
    for (int s = 0; s < PRE; ++s) {
      if (s < steps) {
        uint2 raw[TPW];
#pragma unroll
        for (int i = 0; i < TPW; ++i) raw[i] = pre[i][s];
        step(raw, s);
      }
    }
    for (int s0 = PRE; s0 < steps; s0 += K) {
      uint2 cur[K][TPW];
#pragma unroll
      for (int k = 0; k < K; ++k)
#pragma unroll
        for (int i = 0; i < TPW; ++i) cur[k][i] = nxt[k][i];
      if (s0 + K < steps) load_batch(s0 + K);
#pragma unroll
      for (int k = 0; k < K; ++k)
        if (s0 + k < steps) step(cur[k], s0 + k);
    }
  } else {
    const uint16_t *lhi = reinterpret_cast<const uint16_t *>(lds);
    const int TP = T + 8;
    const uint16_t *llo = lhi + NBT * 16 * TP;
    uint2 nxt[TPW][8];
    auto load_step = [&](int s) {
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int f = ring_frame(r, 32 * s + 8 * sub + q);
#pragma unroll
        for (int i = 0; i < TPW; ++i) nxt[i][q] = dft_load4(r, f, p0[i]);
      }
    };
    for (int s = 0; s < T / 32; ++s) {
      uint2 raw[TPW][8];
#pragma unroll
      for (int i = 0; i < TPW; ++i)
#pragma unroll
        for (int q = 0; q < 8; ++q) raw[i][q] = s == 0 ? pre[i][q] : nxt[i][q];
      if (s + 1 < T / 32) load_step(s + 1);  // next step's frames in flight
      h8 xh[TPW][4];
      bf8 xb[TPW][4];
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          float x[4];
          unpack4(raw[i][q], x);
#pragma unroll
          for (int t = 0; t < 4; ++t) {
            if (MODE == 2)
              xb[i][t][q] = (__bf16)x[t];
            else
              xh[i][t][q] = (_Float16)x[t];
          }
        }
      }
      const int k0 = 32 * s + 8 * sub;
#pragma unroll
      for (int bt = 0; bt < NBT; ++bt) {
        const int row = bt * 16 + (lane & 15);
        const uint4 whi = *reinterpret_cast<const uint4 *>(&lhi[row * TP + k0]);
        if (MODE == 2) {
          bf8 w = __builtin_bit_cast(bf8, whi);
#pragma unroll
          for (int i = 0; i < TPW; ++i)
#pragma unroll
            for (int t = 0; t < 4; ++t)
              acc[i][bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xb[i][t], acc[i][bt][t], 0, 0, 0);
        } else {
          h8 w = __builtin_bit_cast(h8, whi);
#pragma unroll
          for (int i = 0; i < TPW; ++i)
#pragma unroll
            for (int t = 0; t < 4; ++t)
              acc[i][bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(w, xh[i][t], acc[i][bt][t], 0, 0, 0);
          if (MODE == 3) {
            const uint4 wlo = *reinterpret_cast<const uint4 *>(&llo[row * TP + k0]);
            h8 wl = __builtin_bit_cast(h8, wlo);
#pragma unroll
            for (int i = 0; i < TPW; ++i)
#pragma unroll
              for (int t = 0; t < 4; ++t)
                acc[i][bt][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wl, xh[i][t], acc[i][bt][t], 0, 0, 0);
          }
        }
      }
    }
  }
}

// FP32 time DFT of a sequence of rows with a continuous load stream: a ring of K steps
// (4 frames x TPW tiles each) in flight per lane, refilled one step at a time right after
// the step is unpacked, running on into the first K steps of the next row during the last
// K steps of the current one.  With STEPS = T/4 a multiple of K, step s of every row lives
// in slot s % K.  Holds K*TPW*2 registers of loads in flight with no double-buffer copy.
template <int NBT, int TPW, int T, int K> struct DftStream {
  static constexpr int STEPS = T / 4;
  static_assert(STEPS % K == 0, "ring depth must divide the steps of a row");
  uint2 ring[K][TPW];

  __device__ __forceinline__ void load(const DftRing &r, long long p0w, int lane, int s, int slot) {
    const int f = ring_frame(r, 4 * s + (lane >> 4));
    const long long p0 = p0w + 4 * (lane & 15);
#pragma unroll
    for (int i = 0; i < TPW; ++i) ring[slot][i] = dft_load4(r, f, p0 + 64 * i);
  }
  // steps 0 .. K-1 of the first row
  __device__ __forceinline__ void start(const DftRing &r, long long p0w, int lane) {
#pragma unroll
    for (int k = 0; k < K; ++k) load(r, p0w, lane, k, k);
  }
  // D = Wm * X for the row at p0w of window r (operand matrix lw with pitch T + 4 floats);
  // the ring refills from this row, then from the row at p0w_next of window r_next.  On the
  // last row the caller passes any valid row as the next one: those refills are never used,
  // and issuing them anyway keeps every load out of branches (a conditional load makes the
  // steps after it wait for all loads in flight instead of their own).
  __device__ __forceinline__ void row(const DftRing &r, const DftRing &r_next, const float *lw, long long p0w,
                                      long long p0w_next, int lane, f4 (&acc)[TPW][NBT][4]) {
#pragma unroll
    for (int i = 0; i < TPW; ++i)
#pragma unroll
      for (int bt = 0; bt < NBT; ++bt)
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[i][bt][t] = f4{0.f, 0.f, 0.f, 0.f};
    constexpr int TP = T + 4;
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int slot = s % K;
      float x[TPW][4];
#pragma unroll
      for (int i = 0; i < TPW; ++i) unpack4(ring[slot][i], x[i]);
      if (s + K < STEPS)
        load(r, p0w, lane, s + K, slot);
      else  // unconditional: see DftStream::row
        load(r_next, p0w_next, lane, s + K - STEPS, slot);
      __builtin_amdgcn_sched_barrier(0);  // the refill is issued here, not sunk to its use
      const int n = 4 * s + (lane >> 4);
#pragma unroll
      for (int bt = 0; bt < NBT; ++bt) {
        const float w = lw[(bt * 16 + (lane & 15)) * TP + n];
#pragma unroll
        for (int i = 0; i < TPW; ++i)
#pragma unroll
          for (int t = 0; t < 4; ++t)
            acc[i][bt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(w, x[i][t], acc[i][bt][t], 0, 0, 0);
      }
    }
  }
};

// FP32 time DFT of one 16-bin tile with the real-input symmetry folded in (T even):
//   Re X_k = x_0 + (-1)^k x_{T/2} + sum_{p=1}^{T/2-1} (x_p + x_{T-p}) cos(2 pi k p / T)
//   Im X_k =                      - sum_{p=1}^{T/2-1} (x_p - x_{T-p}) sin(2 pi k p / T)
// (cos and sin are even and odd about T/2), which halves the MFMAs of the interleaved
// (cos, -sin) 32-row contraction: per 16-pixel tile 2 (T/8) + 1 MFMAs instead of T/2.
// Step s covers the pairs p = 4 s + g + 1 (g = lane >> 4, the MFMA's K index), loading frames
// p and T - p; the last step's p = T/2 pairs the middle frame with itself under zero weights,
// and x_0, x_{T/2} enter through one more MFMA.  The operand matrix (stage_fold_matrix_rt with
// one tile) holds a cos block and a sin block of 16 rows x PC floats and a 16 x 4 block for
// x_0, x_{T/2}.
// Results: ac[i][t][r] = Re X and as[i][t][r] = Im X of bin 4 g + r for pixel
// p0w + 64 i + 4 (lane & 15) + t.  The ring of K steps (two frames per tile each) runs on into
// the next row as in DftStream.
template <int TPW, int T, int K> struct DftStreamFold {
  static constexpr int STEPS = T / 8;
  static constexpr int PC = T / 2 + 4;  // pitch: 16 rows x 4 K-slots hit 64 distinct banks
  static_assert(STEPS % K == 0, "ring depth must divide the steps of a row");
  uint2 ring[K][TPW][2];

  __device__ __forceinline__ void load(const DftRing &r, long long p0w, int lane, int s, int slot) {
    const int p = 4 * s + (lane >> 4) + 1;
    const int f1 = ring_frame(r, p), f2 = ring_frame(r, T - p);
    const long long p0 = p0w + 4 * (lane & 15);
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      ring[slot][i][0] = dft_load4(r, f1, p0 + 64 * i);
      ring[slot][i][1] = dft_load4(r, f2, p0 + 64 * i);
    }
  }
  __device__ __forceinline__ void start(const DftRing &r, long long p0w, int lane) {
#pragma unroll
    for (int k = 0; k < K; ++k) load(r, p0w, lane, k, k);
  }
  // (refills as in DftStream::row)
  __device__ __forceinline__ void row(const DftRing &r, const DftRing &r_next, const void *lds, long long p0w,
                                      long long p0w_next, int lane, f4 (&ac)[TPW][4], f4 (&as)[TPW][4]) {
    const float *lc = reinterpret_cast<const float *>(lds), *ls = lc + 16 * PC, *lx = ls + 16 * PC;
    const int g = lane >> 4, rw = lane & 15;
#pragma unroll
    for (int i = 0; i < TPW; ++i)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ac[i][t] = f4{0.f, 0.f, 0.f, 0.f};
        as[i][t] = f4{0.f, 0.f, 0.f, 0.f};
      }
    // x_0 (g = 0) and x_{T/2} (g >= 1; zero weights for g >= 2), used after the pair steps
    uint2 xe[TPW];
    {
      const int fe = ring_frame(r, g == 0 ? 0 : T / 2);
      const long long p0 = p0w + 4 * (lane & 15);
#pragma unroll
      for (int i = 0; i < TPW; ++i) xe[i] = dft_load4(r, fe, p0 + 64 * i);
    }
#pragma unroll
    for (int s = 0; s < STEPS; ++s) {
      const int slot = s % K;
      float sa[TPW][4], sb[TPW][4];
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        float x1[4], x2[4];
        unpack4(ring[slot][i][0], x1);
        unpack4(ring[slot][i][1], x2);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          sa[i][t] = x1[t] + x2[t];  // exact: centred 12-bit samples
          sb[i][t] = x1[t] - x2[t];
        }
      }
      if (s + K < STEPS)
        load(r, p0w, lane, s + K, slot);
      else  // unconditional: see DftStream::row
        load(r_next, p0w_next, lane, s + K - STEPS, slot);
      __builtin_amdgcn_sched_barrier(0);  // the refill is issued here, not sunk to its use
      const float wc = lc[rw * PC + 4 * s + g], wsn = ls[rw * PC + 4 * s + g];
#pragma unroll
      for (int i = 0; i < TPW; ++i)
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          ac[i][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wc, sa[i][t], ac[i][t], 0, 0, 0);
          as[i][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wsn, sb[i][t], as[i][t], 0, 0, 0);
        }
    }
    const float wx = lx[rw * 4 + g];
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      float x[4];
      unpack4(xe[i], x);
#pragma unroll
      for (int t = 0; t < 4; ++t) ac[i][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, x[t], ac[i][t], 0, 0, 0);
    }
  }
};

// Runtime-T form of the folded FP32 time DFT (see DftStreamFold) for NBF 16-bin tiles and one
// 64-pixel tile per wave (stand-alone time DFT): steps of four frame pairs in batches of K,
// T % (8 K) == 0.  Operand matrix from stage_fold_matrix_rt (NBF tiles) with pitch T/2 + 4.
template <int NBF, int K = 4>
__device__ __forceinline__ void dft_fold_tile(const DftRing &r, const float *lds, long long p0w, int lane,
                                              f4 (&ac)[NBF][4], f4 (&as)[NBF][4]) {
  const int T = r.T, steps = T / 8, PC = T / 2 + 4;
  const float *lc = lds, *ls = lds + 16 * NBF * PC, *lx = lds + 32 * NBF * PC;
  const int g = lane >> 4, rw = lane & 15;
  const long long p0 = dft_px(r, p0w, lane, 0);
#pragma unroll
  for (int f = 0; f < NBF; ++f)
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      ac[f][t] = f4{0.f, 0.f, 0.f, 0.f};
      as[f][t] = f4{0.f, 0.f, 0.f, 0.f};
    }
  const uint2 xe = dft_load4(r, ring_frame(r, g == 0 ? 0 : T / 2), p0);  // x_0, x_{T/2}
  uint2 ring[K][2];
  auto load = [&](int st, int k) {
    const int p = 4 * st + g + 1;
    ring[k][0] = dft_load4(r, ring_frame(r, p), p0);
    ring[k][1] = dft_load4(r, ring_frame(r, T - p), p0);
  };
#pragma unroll
  for (int k = 0; k < K; ++k) load(k, k);
  // step st from ring slot k, refilling the slot with step st + K when load_next is true_type;
  // the last batch is peeled so that no load sits behind a branch (a conditional load makes
  // every step wait for all loads in flight instead of its own)
  auto step = [&](int st, int k, auto load_next) {
    float x1[4], x2[4], sa[4], sb[4];
    unpack4(ring[k][0], x1);
    unpack4(ring[k][1], x2);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      sa[t] = x1[t] + x2[t];
      sb[t] = x1[t] - x2[t];
    }
    if constexpr (decltype(load_next)::value) {
      load(st + K, k);
      // issue the refill here: left to itself the scheduler sinks all refills of a batch to
      // its end, right before their data is needed
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int f = 0; f < NBF; ++f) {
      const float wc = lc[(16 * f + rw) * PC + 4 * st + g], wsn = ls[(16 * f + rw) * PC + 4 * st + g];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ac[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wc, sa[t], ac[f][t], 0, 0, 0);
        as[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wsn, sb[t], as[f][t], 0, 0, 0);
      }
    }
  };
  int s0 = 0;
  for (; s0 + K < steps; s0 += K) {
#pragma unroll
    for (int k = 0; k < K; ++k) step(s0 + k, k, std::true_type{});
  }
#pragma unroll
  for (int k = 0; k < K; ++k) step(s0 + k, k, std::false_type{});
  float x[4];
  unpack4(xe, x);
#pragma unroll
  for (int f = 0; f < NBF; ++f) {
    const float wx = lx[(16 * f + rw) * 4 + g];
#pragma unroll
    for (int t = 0; t < 4; ++t) ac[f][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(wx, x[t], ac[f][t], 0, 0, 0);
  }
}

// Operand matrix of the folded DFT (DftStreamFold, dft_fold_tile) from the interleaved rows
// wm[2b][n] = cos, wm[2b+1][n] = -sin (b < nbc): for nbf tiles of 16 bins, a cos block
// C[b][p-1] and a sin block S[b][p-1] for the pairs p = 1 .. T/2-1 (zero beyond and for bins
// >= nbc), pitch T/2 + 4, then X[b] = (cos 0, cos(pi k), 0, 0) for x_0, x_{T/2}.
__device__ __forceinline__ void stage_fold_matrix_rt(float *l, const float *wm, int T, int nbf, int nbc, int tid,
                                                     int nthreads) {
  const int PC = T / 2 + 4, rows = 16 * nbf;
  for (int i = tid; i < rows * PC; i += nthreads) {
    const int b = i / PC, j = i % PC;
    const bool in = b < nbc && j < T / 2 - 1;
    l[i] = in ? wm[(2 * b) * T + j + 1] : 0.f;
    l[rows * PC + i] = in ? wm[(2 * b + 1) * T + j + 1] : 0.f;
  }
  for (int i = tid; i < rows * 4; i += nthreads) {
    const int b = i / 4, k = i % 4;
    float w = 0.f;
    if (b < nbc) w = k == 0 ? wm[(2 * b) * T] : k == 1 ? wm[(2 * b) * T + T / 2] : 0.f;
    l[2 * rows * PC + i] = w;
  }
}

__host__ __device__ constexpr size_t dft_fold_lds_bytes(int nbf, int T) {
  return (size_t)nbf * (32 * (T / 2 + 4) + 64) * 4;
}

}  // namespace hfx
