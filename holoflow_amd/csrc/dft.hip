// Time DFT of a uint16 frame window on MFMA (see dft_core.hh), in these forms:
//
//   k_dft          : stand-alone; writes the band planes [planes][H][W] in natural layout
//                    (optionally times the Fresnel lens; negative-band conjugate planes for
//                    the two-sided FFT path).  Used for the per-stage entry point and for
//                    widths the fused kernels do not cover.
//   k_dft_fold     : the same for FP32 with 9..64 bins and T % 32 == 0, with the real-input
//                    symmetry folded into the contraction (about half the MFMAs).
//   k_dft_row      : fused time DFT + forward row FFT.  A persistent workgroup walks a
//                    set of image rows; for each row its waves (64 or 128 pixels each) keep
//                    the band spectra in MFMA registers, push them through LDS into one
//                    padded row per plane, run one wave-level row FFT per plane and write
//                    the result once, column-blocked, for the column pass.  The next row's
//                    first frames are loaded while the current row is transformed.  The
//                    band planes never make an HBM round trip in natural layout.
//   k_front        : the FP32 form of k_dft_row with the waves split into producers (DFT)
//                    and consumers (row FFTs) that overlap across rows (the main path).
//
// Plane q of the output: q < nb -> bin kb+q; q >= nb (two-sided path) -> conj of computed
// bin 2*nb - q, i.e. X[T - (kb + 2*nb - q)] of a real input (fft.cu:163-194 +
// graph_builder.cc:270-287).
#include <type_traits>

#include "dft_core.hh"
#include "hfx_kernels.hh"
#include "wfft.hh"

namespace hfx {

struct DftArgs {
  DftRing r;
  int W;
  int kb, nb, nbc, both;
  float2 *planes;          // k_dft: natural [planes][npix]; k_dft_row: blocked
  const float2 *lx, *ly;   // Fresnel lens vectors or null
  const void *wm_hi, *wm_lo;
  // k_dft_row only
  int H, blk_c, n_planes;
  const float2 *tw;        // W-point FFT twiddles
  // k_front only: the windows of one launch (one plan's consecutive windows, or one window
  // each of a group of plans of one shape): ring, ring length, first frame, planes, lens
  int nwin;
  long long firsts[HFX_MAX_WIN];
  const uint16_t *rings[HFX_MAX_WIN];
  long long ring_frames_w[HFX_MAX_WIN];
  float2 *planes_w[HFX_MAX_WIN];
  const float2 *lx_w[HFX_MAX_WIN], *ly_w[HFX_MAX_WIN];
};

// a wave-uniform value kept in a scalar register from here on (an opaque copy the compiler
// cannot rematerialise by reloading the kernel argument)
__device__ __forceinline__ int pin_sgpr(int v) {
  asm volatile("" : "+s"(v));
  return v;
}

// The band scalars every bin store tests, pinned in scalar registers at the top of a kernel:
// left to the compiler they are reloaded from the kernel arguments inside the stores'
// branches, and each reload's wait (lgkmcnt) also waits for the LDS writes in flight.
struct Band {
  int nb, nbc, kb, both, T;
};
__device__ __forceinline__ Band pinned_band(const DftArgs &a) {
  return {pin_sgpr(a.nb), pin_sgpr(a.nbc), pin_sgpr(a.kb), pin_sgpr(a.both), pin_sgpr(a.r.T)};
}

// Centring offset of computed bin `bin`: 2048 T on a DC bin, else 0.  The DC bin is bin 0, or
// bin T itself, the extra bin the two-sided path computes when its band ends at T (its
// conjugate plane is X[0]).  Added to every bin: a select instead of a branch around the
// additions (adding +0 changes no other value, only the sign of a zero).
__device__ __forceinline__ float dc_offset(int kb, int T, int bin) {
  return kb + bin == 0 || kb + bin == T ? 2048.0f * (float)T : 0.0f;
}

// value of computed bin `bin` for 4 pixels, with the centring offset restored
__device__ __forceinline__ void dft_bin_values(const Band &b, const f4 (&acc)[4], int h, int bin,
                                               float2 (&v)[4]) {
  const float dc = dc_offset(b.kb, b.T, bin);
#pragma unroll
  for (int t = 0; t < 4; ++t) v[t] = make_float2(acc[t][2 * h] + dc, acc[t][2 * h + 1]);
}

// ----------------------------------------------------------------------- stand-alone
// Band bin `bin` of the 4 pixels from p0 into its natural-layout plane (and the conjugate
// plane of the two-sided path), times the lens when there is one.
__device__ __forceinline__ void dft_store_bin(const DftArgs &a, const Band &b, long long p0, int bin,
                                              const float2 (&v)[4], const float2 (&L)[4], bool lens) {
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    int q;
    if (side == 0) {
      if (bin >= b.nb) continue;
      q = bin;
    } else {
      if (!b.both) continue;
      const int j = b.nb - bin;
      if (j < 0 || j >= b.nb) continue;
      q = b.nb + j;
    }
    float2 o[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      float2 c = side ? make_float2(v[t].x, -v[t].y) : v[t];
      o[t] = lens ? cmul(c, L[t]) : c;
    }
    float4 *dst = reinterpret_cast<float4 *>(a.planes + (long long)q * a.r.npix + p0);
    dst[0] = make_float4(o[0].x, o[0].y, o[1].x, o[1].y);
    dst[1] = make_float4(o[2].x, o[2].y, o[3].x, o[3].y);
  }
}

// Persistent: a block stages the operand matrix once (tens of KB, and index arithmetic per
// element) and then walks groups of four 64-pixel tiles blockIdx, blockIdx + grid, ...
template <int NBT, int MODE>
__global__ __launch_bounds__(256) void k_dft(DftArgs a) {
  extern __shared__ float4 lds_dyn[];
  dft_stage_matrix<NBT, MODE>(lds_dyn, a.wm_hi, a.wm_lo, a.r.T, threadIdx.x, 256);
  __syncthreads();
  const Band b = pinned_band(a);
  const long long tiles = (a.r.npix + 63) / 64;
  const bool lens = a.lx != nullptr;
  for (long long t0 = (long long)blockIdx.x * 4; t0 < tiles; t0 += (long long)gridDim.x * 4) {
    // an opaque per-tile copy of the lane index keeps the lane-derived LDS and load addresses of
    // the MFMA loop from being hoisted out of the tile loop (where they would hold registers)
    int lane;
    asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(threadIdx.x & 63));
    // tiles past the end load clamped pixels (dft_px) and store nothing
    const long long p0w = (t0 + (threadIdx.x >> 6)) * 64;
    uint2 pre[1][dft_pre_steps<MODE>()];
    dft_prefetch<MODE, 1>(a.r, p0w, lane, pre);
    f4 acc1[1][NBT][4];
    dft_tile<NBT, MODE, 1>(a.r, lds_dyn, p0w, lane, pre, acc1);
    auto &acc = acc1[0];

    const long long p0 = p0w + 4 * (lane & 15);
    if (p0 >= a.r.npix) continue;
    float2 L[4];
    if (lens) {
      const int y = (int)(p0 / a.W), x = (int)(p0 % a.W);  // W % 4 == 0: same row
      const float2 ly = a.ly[y];
#pragma unroll
      for (int t = 0; t < 4; ++t) L[t] = cmul(a.lx[x + t], ly);
    }
#pragma unroll
    for (int bt = 0; bt < NBT; ++bt) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int bin = bt * 8 + 2 * (lane >> 4) + h;
        if (bin >= b.nbc) continue;
        float2 v[4];
        dft_bin_values(b, acc[bt], h, bin, v);
        dft_store_bin(a, b, p0, bin, v, L, lens);
      }
    }
  }
}

// FP32 stand-alone time DFT with the real-input symmetry folded (dft_fold_tile): NBF tiles of
// 16 bins per wave (all of the band's tiles: every frame sample is loaded by one wave only),
// T % 32 == 0.  Same outputs as k_dft<NBT, 0> up to FP32 rounding, with 2 (T/8) + 1 MFMAs per
// 16 bins and 16 pixels instead of T/2.
// Persistent like k_dft.  The second launch bound (resident blocks per CU) makes the compiler
// keep the accumulators in ordinary registers instead of splitting them off into accumulation
// registers: up to two tiles fit in three waves per SIMD, three and four tiles in two.
constexpr int fold_min_blocks(int nbf) { return nbf <= 2 ? 3 : 2; }
template <int NBF>
__global__ __launch_bounds__(256, fold_min_blocks(NBF)) void k_dft_fold(DftArgs a) {
  extern __shared__ float4 lds_dyn[];
  float *lf = reinterpret_cast<float *>(lds_dyn);
  stage_fold_matrix_rt(lf, reinterpret_cast<const float *>(a.wm_hi), a.r.T, NBF, a.nbc, threadIdx.x, 256);
  __syncthreads();
  const Band b = pinned_band(a);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long tiles = (a.r.npix + 63) / 64;
  const bool lens = a.lx != nullptr;
  for (long long t0 = (long long)blockIdx.x * 4; t0 < tiles; t0 += (long long)gridDim.x * 4) {
    int lane;  // opaque per tile, as in k_dft
    asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(threadIdx.x & 63));
    // tiles past the end load clamped pixels (dft_px) and store nothing
    const long long p0w = (t0 + wave) * 64;
    f4 ac[NBF][4], as[NBF][4];
    dft_fold_tile<NBF>(a.r, lf, p0w, lane, ac, as);

    const long long p0 = p0w + 4 * (lane & 15);
    if (p0 >= a.r.npix) continue;
    float2 L[4];
    if (lens) {
      const int y = (int)(p0 / a.W), x = (int)(p0 % a.W);  // W % 4 == 0: same row
      const float2 ly = a.ly[y];
#pragma unroll
      for (int t = 0; t < 4; ++t) L[t] = cmul(a.lx[x + t], ly);
    }
#pragma unroll
    for (int f = 0; f < NBF; ++f) {
#pragma unroll
      for (int rr = 0; rr < 4; ++rr) {
        const int bin = 16 * f + 4 * (lane >> 4) + rr;
        if (bin >= b.nbc) continue;
        const float dc = dc_offset(b.kb, b.T, bin);
        float2 v[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) v[t] = make_float2(ac[f][t][rr] + dc, as[f][t][rr]);
        dft_store_bin(a, b, p0, bin, v, L, lens);
      }
    }
  }
}

// ----------------------------------------------------------------------- fused with row FFT
// Planes per LDS group: the group's padded rows fill ~127 KB (15 planes at W = 1024), which
// leaves room for the pass twiddles and the operand matrix of the MFMA loop.
template <int W> constexpr int row_group() { return 130560 / (WaveFft<W>::SCR * 8); }

// Waves per row: one per 64 pixels up to 512 wide, one per 128 pixels (two MFMA tiles
// sharing every operand load) beyond, which keeps 8 waves and a 256-register budget.
template <int W> constexpr int row_tpw() { return W > 512 ? 2 : 1; }
template <int W> constexpr int row_threads() { return W / row_tpw<W>(); }

// MULTI: the planes need more than one LDS group (two-sided bands at large widths); a
// separate instantiation so the common single-group kernel carries no spill code.
template <int W, int NBT, int MODE, bool MULTI>
__global__ __launch_bounds__(row_threads<W>()) void k_dft_row(DftArgs a) {
  using F = WaveFft<W>;
  constexpr int TPW = row_tpw<W>();
  constexpr int NT = row_threads<W>();
  constexpr int NW = NT / 64;
  constexpr int E = F::E;
  constexpr int GB = row_group<W>();
  __shared__ float2 scr[GB * F::SCR];  // one padded row per plane of the current group
  __shared__ float2 twl[F::TW];
  extern __shared__ float4 wmat[];     // operand matrix of the MFMA loop

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  dft_stage_matrix<NBT, MODE>(wmat, a.wm_hi, a.wm_lo, a.r.T, tid, NT);
  F::build_twiddles(twl, a.tw, tid, NT);
  __syncthreads();

  const Band b = pinned_band(a);
  const long long plane_elems = (long long)a.H * W;
  const bool lens = a.lx != nullptr;
  const int G = gridDim.x;
  int y = xcd_tile(blockIdx.x, G);
  uint2 pre[TPW][dft_pre_steps<MODE>()];
  if (y < a.H) dft_prefetch<MODE, TPW>(a.r, (long long)y * W + wave * 64 * TPW, lane, pre);

  for (; y < a.H; y += G) {
    // An opaque per-row copy of the lane index keeps the many loop-invariant LDS
    // addresses of the scatter and the transforms out of registers across rows.
    int lane;
    asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(tid & 63));
    f4 acc[TPW][NBT][4];
    dft_tile<NBT, MODE, TPW>(a.r, wmat, (long long)y * W + wave * 64 * TPW, lane, pre, acc);
    // lens factors of the lane's pixels (1 without a lens: the product with (1, 0) is exact,
    // and cheaper than a select)
    auto lens_factors = [&](int i, float2 (&L)[4]) {
      const int x0 = (wave * TPW + i) * 64 + 4 * (lane & 15);
#pragma unroll
      for (int t = 0; t < 4; ++t) L[t] = make_float2(1.f, 0.f);
      if (lens) {
        const float2 ly = a.ly[y];
#pragma unroll
        for (int t = 0; t < 4; ++t) L[t] = cmul(a.lx[x0 + t], ly);
      }
    };
    // single group: the factors are loaded before the next row's prefetch (loaded after it,
    // their wait would also wait for the prefetch: the vector-memory counter drains in order);
    // MULTI loads them per group in the scatter, where registers are scarcer
    float2 Ls[TPW][4];
    if constexpr (!MULTI) {
#pragma unroll
      for (int i = 0; i < TPW; ++i) lens_factors(i, Ls[i]);
    }
    // the next row's first frames load while this row is scattered and transformed
    if (y + G < a.H) dft_prefetch<MODE, TPW>(a.r, (long long)(y + G) * W + wave * 64 * TPW, lane, pre);

    // scatter planes [q0, q0+GB) from the MFMA registers into padded LDS rows
    auto scatter = [&](int q0) {
#pragma unroll
      for (int i = 0; i < TPW; ++i) {
        const int x0 = (wave * TPW + i) * 64 + 4 * (lane & 15);  // first of the lane's 4 pixels
        float2 L[4];
        if constexpr (MULTI)
          lens_factors(i, L);
        else
#pragma unroll
          for (int t = 0; t < 4; ++t) L[t] = Ls[i][t];
#pragma unroll
        for (int bt = 0; bt < NBT; ++bt) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int bin = bt * 8 + 2 * (lane >> 4) + h;
            if (bin >= b.nbc) continue;
            float2 v[4];
            dft_bin_values(b, acc[i][bt], h, bin, v);
#pragma unroll
            for (int side = 0; side < 2; ++side) {
              int q;
              if (side == 0) {
                if (bin >= b.nb) continue;
                q = bin;
              } else {
                if (!b.both) continue;
                const int j = b.nb - bin;
                if (j < 0 || j >= b.nb) continue;
                q = b.nb + j;
              }
              if (q < q0 || q >= q0 + GB) continue;
              float2 *dst = scr + (q - q0) * F::SCR;
#pragma unroll
              for (int t = 0; t < 4; ++t) {
                float2 c = side ? make_float2(v[t].x, -v[t].y) : v[t];
                dst[F::pad(x0 + t)] = cmul(c, L[t]);
              }
            }
          }
        }
      }
    };
    // one wave per plane row of the group: forward row FFT, then the column-blocked store
    auto transform = [&](int q0) {
      const int q1 = min(q0 + GB, a.n_planes);
      const long long row_off = blk_lane_offset(lane, a.H, a.blk_c) + (long long)y * a.blk_c;
      for (int q = q0 + wave; q < q1; q += NW) {
        float2 *ws = scr + (q - q0) * F::SCR;
        float2 v[E];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = ws[F::pad(lane + 64 * m)];
        F::template run<false>(v, ws, twl, lane);
        float2 *dst = a.planes + (long long)q * plane_elems + row_off;
#pragma unroll
        for (int m = 0; m < E; ++m) dst[(long long)m * 64 * a.H] = v[m];
      }
    };
    if constexpr (!MULTI) {
      // every plane fits one group, so the MFMA accumulators are dead before the
      // transforms need their registers
      scatter(0);
      __syncthreads();
      transform(0);
      __syncthreads();
    } else {
      for (int q0 = 0; q0 < a.n_planes; q0 += GB) {
        scatter(q0);
        __syncthreads();
        transform(q0);
        __syncthreads();
      }
    }
  }
}

// ----------------------------------------------------------------------- pipelined front
// Same result as k_dft_row for a single LDS group of planes (up to FP32 rounding where the
// DFT is folded), with the waves split by role: W/128 producer waves (8 at W = 1024) run the
// MFMA time DFT of the block's next row into registers while four consumer waves run the row
// FFTs of the current row's planes out of LDS and store them; then the producers scatter the
// next row into LDS.  HBM loads and MFMAs of one row overlap the LDS/VALU work and the stores
// of the previous one.
// Producers own 128 pixels each (two 64-pixel MFMA tiles) and stream their frame loads
// through a register ring that runs on across rows: DftStream (interleaved cos/-sin rows, up
// to 8 bins) or DftStreamFold (9..16 bins, real-input symmetry folded).  FP32 only, for T in
// {32, 64, 128}.  HFX_FRONT_CW / HFX_FRONT_RING override the consumer-wave count and the ring
// depth at compile time (tuning builds; the defaults are the measured best).
template <int W> constexpr int front_pw() { return W / 128; }
#ifndef HFX_FRONT_CW
#define HFX_FRONT_CW 4
#endif
template <int W> constexpr int front_cw() { return HFX_FRONT_CW; }
template <int W> constexpr int front_threads() { return (front_pw<W>() + front_cw<W>()) * 64; }
#ifndef HFX_FRONT_RING
#define HFX_FRONT_RING 8
#endif
constexpr int front_ring(int T) { return T / 4 < HFX_FRONT_RING ? T / 4 : HFX_FRONT_RING; }

template <int W, int NBT, int MODE, int T>
__global__ __launch_bounds__(front_threads<W>()) void k_front(DftArgs a) {
  static_assert(MODE == 0, "the pipelined front kernel runs the FP32 time DFT");
  using F = WaveFft<W>;
  constexpr int TPW = 2;
  constexpr int NPW = front_pw<W>();
  constexpr int NT = front_threads<W>();
  constexpr int E = F::E;
  constexpr int GB = row_group<W>();
  __shared__ float2 scr[GB * F::SCR];  // one padded row per plane
  __shared__ float2 twl[F::TW];
  extern __shared__ float4 wmat[];     // operand matrix of the MFMA loop

  // one 16-bin tile (NBT == 2: up to 16 bins as 32 interleaved cos/-sin rows): the folded
  // contraction of DftStreamFold, half the MFMAs
  constexpr bool FOLD = NBT == 2;
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const bool producer = wave < NPW;    // wave-uniform role
  const int rw = producer ? wave : wave - NPW;
  if constexpr (FOLD)
    stage_fold_matrix_rt(reinterpret_cast<float *>(wmat), reinterpret_cast<const float *>(a.wm_hi), T, 1, a.nbc,
                         tid, NT);
  else
    dft_stage_matrix<NBT, MODE>(wmat, a.wm_hi, a.wm_lo, a.r.T, tid, NT);
  F::build_twiddles(twl, a.tw, tid, NT);
  __syncthreads();

  const long long plane_elems = (long long)a.H * W;
  const bool lens = a.lx_w[0] != nullptr;
  const int G = gridDim.x;
  const int y0 = xcd_tile(blockIdx.x, G);
  const int nrw = y0 < a.H ? (a.H - 1 - y0) / G + 1 : 0;  // rows of one window
  const int nr = nrw * a.nwin;                             // rows of all windows, window-major

  f4 acc[TPW][NBT][4];        // interleaved rows (!FOLD)
  f4 fre[TPW][4], fim[TPW][4];  // Re and Im of bins 4 g + r (FOLD)
  std::conditional_t<FOLD, DftStreamFold<TPW, T, 4>, DftStream<NBT, TPW, T, front_ring(T)>> stream;
  auto px0 = [&](int y) { return (long long)y * W + rw * 64 * TPW; };
  auto row_y = [&](int i) { return y0 + (nrw ? i % nrw : 0) * G; };
  auto win_of = [&](int i) { return nrw ? i / nrw : 0; };
  auto ring_of = [&](int i) {
    const int w = win_of(i);
    DftRing r = a.r;
    r.ring = a.rings[w];
    r.ring_frames = a.ring_frames_w[w];
    r.first = a.firsts[w];
    return r;
  };

  // producers: the band spectra of row i for their 128 pixels
  auto dft_row = [&](int i, int lane) {
    // the last row's refills (never used) all read the first pixels of the window's frames:
    // every wave of every CU loads the same few KB, which stay in L2 instead of costing HBM reads
    const bool next = i + 1 < nr;
    const int j = next ? i + 1 : i;
    const long long pn = next ? px0(row_y(j)) : 0;
    if constexpr (FOLD)
      stream.row(ring_of(i), ring_of(j), wmat, px0(row_y(i)), pn, lane, fre, fim);
    else
      stream.row(ring_of(i), ring_of(j), reinterpret_cast<const float *>(wmat), px0(row_y(i)), pn, lane, acc);
  };
  const Band b = pinned_band(a);
  // producers: MFMA registers of row i -> padded LDS rows (its window's lens for Fresnel)
  auto scatter = [&](int i, int lane) {
    const int y = row_y(i);
    const float2 *lxw = a.lx_w[win_of(i)], *lyw = a.ly_w[win_of(i)];
    // band bin `bin` of 4 pixels from x0 into the LDS plane rows (and its conjugate plane on
    // the two-sided path), lens applied.  The product's rounding is pinned (cmul_rn): the
    // first row is scattered by the copy of this code before the row loop, the others by the
    // copy inside it, and a window's first row may fall in either (a grouped launch), so both
    // copies must round alike.
    auto put = [&](int x0, int bin, const float2 (&v)[4], const float2 (&L)[4]) {
#pragma unroll
      for (int side = 0; side < 2; ++side) {
        int q;
        if (side == 0) {
          if (bin >= b.nb) continue;
          q = bin;
        } else {
          if (!b.both) continue;
          const int j = b.nb - bin;
          if (j < 0 || j >= b.nb) continue;
          q = b.nb + j;
        }
        float2 *dst = scr + q * F::SCR;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          float2 c = side ? make_float2(v[t].x, -v[t].y) : v[t];
          dst[F::pad(x0 + t)] = cmul_rn(c, L[t]);
        }
      }
    };
#pragma unroll
    for (int i = 0; i < TPW; ++i) {
      const int x0 = (rw * TPW + i) * 64 + 4 * (lane & 15);
      // lens factors; 1 without a lens (the product with (1, 0) is exact, and cheaper than a select)
      float2 L[4] = {make_float2(1.f, 0.f), make_float2(1.f, 0.f), make_float2(1.f, 0.f), make_float2(1.f, 0.f)};
      if (lens) {
        const float2 ly = lyw[y];
#pragma unroll
        for (int t = 0; t < 4; ++t) L[t] = cmul(lxw[x0 + t], ly);
      }
      if constexpr (FOLD) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          // bins past the computed ones fail put's own tests (nbc = nb, or nb + 1 on the
          // two-sided path, whose conjugate planes use bins 1..nb)
          const int bin = 4 * (lane >> 4) + r;
          // the centring offset restored on the DC bin: a select, not a branch
          const float dc = dc_offset(b.kb, T, bin);
          float2 v[4];
#pragma unroll
          for (int t = 0; t < 4; ++t) v[t] = make_float2(fre[i][t][r] + dc, fim[i][t][r]);
          put(x0, bin, v, L);
        }
      } else {
#pragma unroll
        for (int bt = 0; bt < NBT; ++bt) {
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int bin = bt * 8 + 2 * (lane >> 4) + h;
            if (bin >= b.nbc) continue;
            float2 v[4];
            dft_bin_values(b, acc[i][bt], h, bin, v);
            put(x0, bin, v, L);
          }
        }
      }
    }
  };

  // consumers: one wave per plane row: forward row FFT, then the column-blocked store
  auto transform = [&](int i, int lane) {
    const int y = row_y(i);
    float2 *planes = a.planes_w[win_of(i)];
    const long long row_off = blk_lane_offset(lane, a.H, a.blk_c) + (long long)y * a.blk_c;
    for (int q = rw; q < a.n_planes; q += front_cw<W>()) {
      float2 *ws = scr + q * F::SCR;
      float2 v[E];
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = ws[F::pad(lane + 64 * m)];
      F::template run<false>(v, ws, twl, lane);
      float2 *dst = planes + (long long)q * plane_elems + row_off;
#pragma unroll
      for (int m = 0; m < E; ++m) dst[(long long)m * 64 * a.H] = v[m];
    }
  };

  // The two roles run separate loops (wave-uniform branch) that meet at the same number of
  // workgroup barriers: 1 + 2 per row.  Keeping them apart keeps the producers' MFMA
  // accumulators out of the consumers' register allocation.
  if (producer) {
    const int lane0 = tid & 63;
    if (nr > 0) {
      stream.start(ring_of(0), px0(y0), lane0);
      dft_row(0, lane0);
      scatter(0, lane0);
    }
    __syncthreads();
    for (int i = 0; i < nr; ++i) {
      int lane;
      asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(tid & 63));
      if (i + 1 < nr) dft_row(i + 1, lane);
      __syncthreads();  // consumers are done with row i's planes
      if (i + 1 < nr) scatter(i + 1, lane);
      __syncthreads();  // row i + 1's planes are in LDS
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the last row's unused refills land first
  } else {
    __syncthreads();
    for (int i = 0; i < nr; ++i) {
      // opaque per-row lane index: keeps loop-invariant LDS addresses out of registers
      int lane;
      asm volatile("v_mov_b32 %0, %1" : "=v"(lane) : "v"(tid & 63));
      transform(i, lane);
      __syncthreads();
      __syncthreads();
    }
  }
}

// ----------------------------------------------------------------------- launch
static DftArgs make_args(const DftLaunch &L) {
  DftArgs a{};
  a.r.ring = L.ring;
  a.r.ring_frames = L.ring_frames;
  a.r.first = L.first;
  a.r.npix = L.npix;
  a.r.T = L.T;
  a.W = L.W;
  a.kb = L.kb;
  a.nb = L.nb;
  a.nbc = L.nbc;
  a.both = L.both;
  a.planes = L.planes;
  a.lx = L.lx;
  a.ly = L.ly;
  switch (L.precision) {
    case 0: a.wm_hi = L.wm_f32; break;
    case 2: a.wm_hi = L.wm_bf16; break;
    default: a.wm_hi = L.wm_f16hi; break;
  }
  a.wm_lo = L.wm_f16lo;
  a.nwin = L.nwin > 0 ? L.nwin : 1;
  for (int w = 0; w < HFX_MAX_WIN; ++w) {
    const bool own = w < a.nwin && L.rings[w] != nullptr;  // a window of another plan of the group
    a.firsts[w] = w == 0 ? L.first : w < a.nwin ? L.firsts[w] : L.first;
    a.rings[w] = own ? L.rings[w] : L.ring;
    a.ring_frames_w[w] = own ? L.ring_frames_w[w] : L.ring_frames;
    a.planes_w[w] = own ? L.planes_w[w] : L.planes + (w < a.nwin ? w : 0) * L.planes_stride;
    a.lx_w[w] = own ? L.lx_w[w] : L.lx;
    a.ly_w[w] = own ? L.ly_w[w] : L.ly;
  }
  return a;
}

// Grid of a persistent stand-alone DFT kernel: as many blocks as are resident at once (the
// occupancy calculator's count per CU times the CUs), at most one per work item.
static unsigned persistent_grid(const void *kernel, size_t lds, long long items) {
  int dev = 0, cus = 256, per_cu = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 256, lds) != hipSuccess || per_cu < 1) per_cu = 1;
  const long long g = (long long)cus * per_cu;
  return (unsigned)(items < g ? (items > 0 ? items : 1) : g);
}

template <int NBT, int MODE> static int launch_plain(const DftArgs &a, hipStream_t st) {
  const long long waves = (a.r.npix + 63) / 64;
  const size_t lds = dft_matrix_lds_bytes(NBT, MODE, a.r.T);
  const unsigned grid = persistent_grid(reinterpret_cast<const void *>(&k_dft<NBT, MODE>), lds, (waves + 3) / 4);
  hipLaunchKernelGGL((k_dft<NBT, MODE>), dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int NBT> static int launch_plain_mode(const DftArgs &a, int mode, hipStream_t st) {
  switch (mode) {
    case 0: return launch_plain<NBT, 0>(a, st);
    case 1: return launch_plain<NBT, 1>(a, st);
    case 2: return launch_plain<NBT, 2>(a, st);
    case 3: return launch_plain<NBT, 3>(a, st);
  }
  return -2;
}

template <int NBF> static int launch_fold(const DftArgs &a, hipStream_t st) {
  const long long waves = (a.r.npix + 63) / 64;
  const size_t lds = dft_fold_lds_bytes(NBF, a.r.T);
  const unsigned grid = persistent_grid(reinterpret_cast<const void *>(&k_dft_fold<NBF>), lds, (waves + 3) / 4);
  hipLaunchKernelGGL((k_dft_fold<NBF>), dim3(grid), dim3(256), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_time_dft(const DftLaunch &L, hipStream_t st) {
  const DftArgs a = make_args(L);
  // FP32 with more than one 16-row tile of interleaved rows: the folded contraction
  if (L.precision == 0 && L.T % 32 == 0 && L.nbc > 8 && L.nbc <= 64) {
    switch ((L.nbc + 15) / 16) {
      case 1: return launch_fold<1>(a, st);
      case 2: return launch_fold<2>(a, st);
      case 3: return launch_fold<3>(a, st);
      case 4: return launch_fold<4>(a, st);
    }
  }
  switch ((2 * L.nbc + 15) / 16) {
    case 1: return launch_plain_mode<1>(a, L.precision, st);
    case 2: return launch_plain_mode<2>(a, L.precision, st);
    case 3: return launch_plain_mode<3>(a, L.precision, st);
    case 4: return launch_plain_mode<4>(a, L.precision, st);
    case 5: return launch_plain_mode<5>(a, L.precision, st);
    case 6: return launch_plain_mode<6>(a, L.precision, st);
    case 7: return launch_plain_mode<7>(a, L.precision, st);
    case 8: return launch_plain_mode<8>(a, L.precision, st);
  }
  return -2;
}

bool dft_row_supported(int W, int nbc, int T) {
  return (W == 512 || W == 768 || W == 1024) && (2 * nbc + 15) / 16 <= 2 && T <= 128;
}

bool dft_front_pipelined(int W, int nbc, int T, int n_planes, int precision) {
  if (!dft_row_supported(W, nbc, T) || precision != 0 || !(T == 32 || T == 64 || T == 128)) return false;
  const int group = W == 512 ? row_group<512>() : W == 768 ? row_group<768>() : row_group<1024>();
  return n_planes <= group;
}

template <int W, int NBT, int MODE> static int launch_row(const DftArgs &a, hipStream_t st) {
  const size_t lds = dft_matrix_lds_bytes(NBT, MODE, a.r.T);
  // persistent grid: one workgroup per CU (the LDS footprint allows no more), each
  // walking rows blockIdx, blockIdx + grid, ...
  int dev = 0, cus = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const unsigned grid = (unsigned)(a.H < cus ? a.H : cus);
  const int T = a.r.T;
  if (MODE == 0 && a.n_planes <= row_group<W>() && (T == 32 || T == 64 || T == 128)) {
    if (T == 32)
      hipLaunchKernelGGL((k_front<W, NBT, 0, 32>), dim3(grid), dim3(front_threads<W>()), lds, st, a);
    else if (T == 64)
      hipLaunchKernelGGL((k_front<W, NBT, 0, 64>), dim3(grid), dim3(front_threads<W>()), lds, st, a);
    else
      hipLaunchKernelGGL((k_front<W, NBT, 0, 128>), dim3(grid), dim3(front_threads<W>()), lds, st, a);
  } else if (a.n_planes <= row_group<W>())
    hipLaunchKernelGGL((k_dft_row<W, NBT, MODE, false>), dim3(grid), dim3(row_threads<W>()), lds, st, a);
  else
    hipLaunchKernelGGL((k_dft_row<W, NBT, MODE, true>), dim3(grid), dim3(row_threads<W>()), lds, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int W, int NBT> static int launch_row_mode(const DftArgs &a, int mode, hipStream_t st) {
  switch (mode) {
    case 0: return launch_row<W, NBT, 0>(a, st);
    case 1: return launch_row<W, NBT, 1>(a, st);
    case 2: return launch_row<W, NBT, 2>(a, st);
    case 3: return launch_row<W, NBT, 3>(a, st);
  }
  return -2;
}

template <int W> static int launch_row_w(const DftArgs &a, int nbt, int mode, hipStream_t st) {
  if (nbt == 1) return launch_row_mode<W, 1>(a, mode, st);
  if (nbt == 2) return launch_row_mode<W, 2>(a, mode, st);
  return -2;
}

int launch_time_dft_row(const DftLaunch &L, int H, int blk_c, int n_planes, const float2 *tw,
                        hipStream_t st) {
  if (!dft_row_supported(L.W, L.nbc, L.T)) return -2;
  if (L.nwin > 1 && (L.nwin > HFX_MAX_WIN || !dft_front_pipelined(L.W, L.nbc, L.T, n_planes, L.precision)))
    return -6;
  DftArgs a = make_args(L);
  a.H = H;
  a.blk_c = blk_c;
  a.n_planes = n_planes;
  a.tw = tw;
  const int nbt = (2 * L.nbc + 15) / 16;
  switch (L.W) {
    case 512: return launch_row_w<512>(a, nbt, L.precision, st);
    case 768: return launch_row_w<768>(a, nbt, L.precision, st);
    case 1024: return launch_row_w<1024>(a, nbt, L.precision, st);
  }
  return -2;
}

}  // namespace hfx
