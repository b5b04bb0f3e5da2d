// Row and column passes built on the wave-level FFT (wfft.hh) for transform lengths
// that are multiples of 64 (256, 512, 768, 1024, 2048).  Same kernel kinds and layouts
// as the workgroup-wide passes in fft2d.hip, which remain for short lengths:
//
//   row pass    : one wave per image row; a workgroup holds 4 consecutive rows (tiles are
//                 spread over the XCDs in contiguous runs).  BLK_TO_MOM waves loop over
//                 all planes of their row, prefetching the next plane into registers.
//   column pass : one wave per column of a column block (C <= 4 columns).  BLK_TO_MOM
//                 workgroups hold S plane subsets x 4 columns (S = 2, or 1 for 2048 rows);
//                 the subsets' partial moment sums are added in a fixed order at the
//                 end, so results are deterministic.
#include <type_traits>

#include "hfx_kernels.hh"
#include "wfft.hh"

namespace hfx {

namespace {

// BLK_TO_MOM5: moments plus the two band-split M0 images (5 sums instead of 3)
enum Kind { NAT_TO_BLK = 0, BLK_TO_MOM = 1, BLK_MID = 2, BLK_TO_NAT = 3, BLK_TO_MOM5 = 4 };
constexpr bool is_mom(int k) { return k == BLK_TO_MOM || k == BLK_TO_MOM5; }
// Transform direction each kind is compiled for: the row pass out of natural data is
// forward (it feeds the column pass), the row passes that read the blocked spectrum back
// (angular spectrum) are inverse, and every column pass is forward (BLK_MID runs forward,
// transfer function, inverse).  The launchers reject a PassArgs::inv that disagrees.
constexpr bool row_inv(int k) { return k != NAT_TO_BLK; }

template <int E, int NACC> struct Moments {
  float s[NACC][E];
  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int k = 0; k < NACC; ++k)
#pragma unroll
      for (int m = 0; m < E; ++m) s[k][m] = 0.f;
  }
  // f, f2: the plane's M1/M2 weights, read by the caller at the top of its plane iteration
  // (wave-uniform scalar loads) so that their latency is not paid after the transform
  __device__ __forceinline__ void add(const float2 (&v)[E], float f, float f2, bool lo, bool power2) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float a = sq_mag(v[m]);
      if (!power2) a = mag_sqrt(a);
      s[0][m] += a;
      s[1][m] += __fmul_rn(f, a);
      s[2][m] += __fmul_rn(f2, a);
      if constexpr (NACC == 5) {
        if (lo)
          s[3][m] += a;
        else
          s[4][m] += a;
      }
    }
  }
  // the three moment sums, plus sum 3 (NACC == 4) when HI: for the row pass's band split
  template <bool HI>
  __device__ __forceinline__ void add_split(const float2 (&v)[E], float f, float f2, bool power2) {
#pragma unroll
    for (int m = 0; m < E; ++m) {
      float a = sq_mag(v[m]);
      if (!power2) a = mag_sqrt(a);
      s[0][m] += a;
      s[1][m] += __fmul_rn(f, a);
      s[2][m] += __fmul_rn(f2, a);
      if constexpr (HI) s[3][m] += a;
    }
  }
};

// The pointers of this block's window: win[blockIdx.z] of a grouped launch (G), else the
// pass's own.  Grouped launches are separate instantiations: reading the pointers through the
// window table changes how the compiler schedules the loads of the single-window kernels
// (measured 685 -> 1270 us for the 2048-point moment row pass).
template <bool G> __device__ __forceinline__ PassWin pass_window(const PassArgs &a) {
  if constexpr (G) {
    return a.win[blockIdx.z];
  } else {
    PassWin w;
    w.in = a.in;
    w.out = a.out;
    w.mid_tf = a.mid_tf;
#pragma unroll
    for (int k = 0; k < 5; ++k) w.mom_out[k] = a.mom.out[k];
    return w;
  }
}

// offset of pixel (y, x) in a moment image (fftshift applied on store when asked)
__device__ __forceinline__ long long moment_index(const PassArgs &a, int y, int x) {
  int ys = y, xs = x;  // 0 <= y < H, 0 <= x < W
  if (a.mom.shift) {   // (y + H - H/2) mod H, without a division
    ys = y + (a.H - a.H / 2);
    if (ys >= a.H) ys -= a.H;
    xs = x + (a.W - a.W / 2);
    if (xs >= a.W) xs -= a.W;
  }
  return (long long)ys * a.W + xs;
}

__device__ __forceinline__ void store_moment(const PassArgs &a, float *const (&mo)[5], int nacc, int y, int x,
                                             const float (&val)[5]) {
  const long long o = moment_index(a, y, x);
  const float n_all = (float)a.planes;
  mo[0][o] = val[0] / n_all;
  mo[1][o] = val[1] / n_all;
  mo[2][o] = val[2] / n_all;
  if (nacc == 5) {
    if (mo[3]) mo[3][o] = val[3] / (float)a.mom.split;
    if (mo[4]) mo[4][o] = val[4] / (float)(a.planes - a.mom.split);
  }
}

}  // namespace

// ------------------------------------------------------------------------- row pass
// Rows (waves) per workgroup.  Each wave owns a padded scratch row in LDS, and the workgroup
// shares one twiddle table, so the LDS footprint sets the occupancy: at 2048 points four
// waves (86 KB) leave one workgroup, one wave per SIMD, per CU, while eight waves in one
// workgroup (155 KB) give two waves per SIMD to hide the memory latency.  The moment kinds
// keep four: their 5*E sums leave no register room for a second wave per SIMD.
constexpr int wrow_waves(int n, int kind) { return n >= 2048 && !is_mom(kind) ? 8 : 4; }

// PLAIN (NAT_TO_BLK only): the input is a complex plane of the output's size with no lens,
// so a row loads from one base address with no per-element bounds checks or index products.
template <int N, int KIND, bool G = false, bool PLAIN = false>
__global__ __launch_bounds__(64 * wrow_waves(N, KIND)) void k_wrow(PassArgs a) {
  using F = WaveFft<N>;
  constexpr int E = F::E;
  constexpr int RW = wrow_waves(N, KIND);
  __shared__ float2 scr[RW * F::SCR];
  __shared__ float2 twl[F::TW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  F::build_twiddles(twl, a.tw, tid, 64 * RW);
  __syncthreads();

  const int y = xcd_tile(blockIdx.x, gridDim.x) * RW + wave;
  if (y >= a.H) return;
  const PassWin pw = pass_window<G>(a);
  const int C = a.blk_c;
  const long long plane_elems = (long long)a.H * a.W;
  const long long boff = blk_lane_offset(lane, a.H, C) + (long long)y * C;  // blocked (y, lane)
  float2 *ws = scr + wave * F::SCR;

  auto load = [&](float2 (&v)[E], long long p) {
    if constexpr (PLAIN) {
      const float2 *row = reinterpret_cast<const float2 *>(pw.in) + p * plane_elems + (long long)y * N;
#pragma unroll
      for (int m = 0; m < E; ++m) v[m] = row[lane + 64 * m];
      return;
    }
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = lane + 64 * m;
      if (KIND == NAT_TO_BLK) {
        const int yi = y - a.off_y, xi = x - a.off_x;
        float2 t = make_float2(0.f, 0.f);
        if (yi >= 0 && yi < a.in_h && xi >= 0 && xi < a.in_w) {
          const long long o = p * (long long)a.in_h * a.in_w + (long long)yi * a.in_w + xi;
          if (a.in_mode == IN_NAT_C)
            t = reinterpret_cast<const float2 *>(pw.in)[o];
          else
            t = make_float2(reinterpret_cast<const float *>(pw.in)[o], 0.f);
          if (a.pre_lx) t = cmul(t, cmul(a.pre_lx[xi], a.pre_ly[yi]));
        }
        v[m] = t;
      } else {
        v[m] = reinterpret_cast<const float2 *>(pw.in)[p * plane_elems + boff + (long long)m * 64 * a.H];
      }
    }
  };

  if constexpr (is_mom(KIND)) {
    // Sums over the planes of the row.  The band-split kind also needs M0 over the planes below
    // the split and over those above it: the planes below run first with the three sums, M0
    // low is stored from sum 0 at the split (the same additions in the same order as a sum of
    // its own), then the planes above run with a fourth sum for M0 high.  Fewer sums are live
    // at a time, and at 2048 points fewer of them spill into accumulation registers.
    constexpr bool BAND = KIND == BLK_TO_MOM5;
    Moments<E, BAND ? 4 : 3> acc;
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
      for (int m = 0; m < E; ++m) acc.s[k][m] = 0.f;
    // the next plane's row is loaded once pass 1 has moved this one into the scratch, so its
    // latency overlaps the remaining passes; the load is unconditional (the last one reloads
    // the last plane, unused), which keeps it out of a branch
    float2 v[E];
    load(v, 0);
    auto run_planes = [&](int p0, int p1, auto hi) {
      for (int p = p0; p < p1; ++p) {
        const float f = a.mom.f[p], f2 = a.mom.f2[p];
        F::template stage_in<row_inv(KIND)>(v, ws, lane);
        load(v, p + 1 < a.planes ? p + 1 : p);
        float2 o[E];
        F::template stage_out<row_inv(KIND)>(o, ws, twl, lane);
        acc.template add_split<decltype(hi)::value>(o, f, f2, a.mom.power2);
      }
    };
    if constexpr (BAND) {
      run_planes(0, a.mom.split, std::false_type{});
      if (pw.mom_out[3]) {
#pragma unroll
        for (int m = 0; m < E; ++m)
          pw.mom_out[3][moment_index(a, y, lane + 64 * m)] = acc.s[0][m] / (float)a.mom.split;
      }
#pragma unroll
      for (int m = 0; m < E; ++m) acc.s[3][m] = 0.f;
      run_planes(a.mom.split, a.planes, std::true_type{});
    } else {
      run_planes(0, a.planes, std::false_type{});
    }
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the unused last refill lands first
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const long long o = moment_index(a, y, lane + 64 * m);
      const float n_all = (float)a.planes;
      pw.mom_out[0][o] = acc.s[0][m] / n_all;
      pw.mom_out[1][o] = acc.s[1][m] / n_all;
      pw.mom_out[2][o] = acc.s[2][m] / n_all;
      if constexpr (BAND) {
        if (pw.mom_out[4]) pw.mom_out[4][o] = acc.s[3][m] / (float)(a.planes - a.mom.split);
      }
    }
  } else {
    const long long p = blockIdx.y;
    float2 v[E];
    load(v, p);
    F::template run<row_inv(KIND)>(v, ws, twl, lane);
    // blocked store: element m of the lane's row at dst + m * mstep, the step a wave-uniform scalar
    float2 *dst = reinterpret_cast<float2 *>(pw.out) + p * plane_elems + boff;
    const long long mstep = 64LL * a.H;
#pragma unroll
    for (int m = 0; m < E; ++m) {
      const int x = lane + 64 * m;
      float2 t = v[m];
      if (KIND == NAT_TO_BLK) {
        dst[m * mstep] = t;
      } else {
        const long long o = p * plane_elems + (long long)y * a.W + x;
        if (a.out_mode == OUT_NAT_MAG) {
          reinterpret_cast<float *>(pw.out)[o] = sqrtf(t.x * t.x + t.y * t.y);
        } else if (a.out_mode == OUT_NAT_RE) {
          reinterpret_cast<float *>(pw.out)[o] = t.x;
        } else {
          if (a.post_lx) t = cmul(t, cmul(a.post_lx[x], a.post_ly[y]));
          reinterpret_cast<float2 *>(pw.out)[o] = t;
        }
      }
    }
  }
}

// ------------------------------------------------------------------------- column pass
// Plane subsets per column tile (waves = 4 * subsets).  A wave keeps its column (E
// values) and 5*E moment sums in registers.  Three subsets (12 waves, 3 per SIMD) without
// a register prefetch measured faster at 1024 rows than two subsets with the next plane
// prefetched into registers; 2048-row columns leave room for one subset only.
// The five-sum kind (band split) keeps 5*E sums per lane: two subsets at 1024 rows (8 waves,
// 256 registers each) instead of three, which spilled.
constexpr int wcol_subsets(int n, int kind) { return n >= 2048 ? 1 : kind == BLK_TO_MOM5 ? 2 : 3; }
constexpr bool wcol_prefetch(int n, int kind) { return n >= 2048 ? false : wcol_subsets(n, kind) < 3; }
// The plane-parallel column kinds take one plane per block: a register prefetch of the next
// plane's column costs waves or spills; at 2048 rows two waves per SIMD hide the latency
// instead (wcol_tiles).
// Column tiles per workgroup of the plane-parallel kinds: two at 2048 rows (eight waves share
// one twiddle table: 155 KB of LDS, two waves per SIMD instead of one), else one.
constexpr int wcol_tiles(int n, int kind) { return n >= 2048 && !is_mom(kind) ? 2 : 1; }
constexpr int wcol_threads(int n, int kind) {
  return is_mom(kind) ? 256 * wcol_subsets(n, kind) : 256 * wcol_tiles(n, kind);
}

template <int N, int KIND, bool G = false>
__global__ __launch_bounds__(wcol_threads(N, KIND)) void k_wcol(PassArgs a) {
  using F = WaveFft<N>;
  constexpr int E = F::E;
  constexpr int S = is_mom(KIND) ? wcol_subsets(N, KIND) : 1;
  constexpr int TL = wcol_tiles(N, KIND);
  constexpr int WAVES = 4 * S * TL;
  __shared__ float2 scr[WAVES * F::SCR];
  __shared__ float2 twl[F::TW];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  F::build_twiddles(twl, a.tw, tid, 64 * WAVES);
  __syncthreads();

  const int C = a.blk_c;
  // wave-uniform to the compiler too: plane indices, weights and tile offsets become scalars
  const int c = __builtin_amdgcn_readfirstlane(wave & 3), sub = __builtin_amdgcn_readfirstlane(wave >> 2);
  // column tiles in contiguous runs per XCD: the moment images are written column-wise, so
  // the 8 tiles that share an output cache line (32 floats) then share an L2 that merges
  // their 16-byte pieces before they go out
  // plane-parallel kinds with TL tiles: wave >> 2 picks the tile (sub is 0 for them)
  // BLK_MID multiplies every plane's column by the same transfer-function tile, so its blocks
  // run plane-fastest (block = tile * planes + plane, in contiguous runs per XCD): the planes
  // of one tile are in flight together on one XCD and read the table tile from its L2,
  // instead of the whole table streaming from memory once per plane.
  int bt, bp;
  if constexpr (KIND == BLK_MID) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    bt = b / a.planes;
    bp = b - bt * a.planes;
  } else {
    bt = xcd_tile(blockIdx.x, gridDim.x);
    bp = blockIdx.y;
  }
  const int bx = bt * TL + (TL > 1 ? (int)(wave >> 2) : 0);
  const int x = bx * C + c;
  const bool active = c < C && bx * C < a.W;
  const long long plane_elems = (long long)a.H * a.W;
  const long long tile = (long long)bx * a.H * C;  // blocked offset of the tile
  // element (lane + 64 m) of the column: tile + lane * C + c + m * col_step, the per-m step a
  // wave-uniform scalar (no per-element 64-bit products kept live across the transforms)
  const long long col_step = 64LL * C;
  const PassWin pw = pass_window<G>(a);
  float2 *ws = scr + wave * F::SCR;

  auto load = [&](float2 (&v)[E], long long p) {
#pragma unroll
    for (int m = 0; m < E; ++m)
      v[m] = reinterpret_cast<const float2 *>(pw.in)[p * plane_elems + tile + lane * C + c + m * col_step];
  };

  if constexpr (is_mom(KIND)) {
    constexpr int NACC = KIND == BLK_TO_MOM5 ? 5 : 3;
    Moments<E, NACC> acc;
    acc.zero();
    if (active && wcol_prefetch(N, KIND)) {
      float2 v[E], nxt[E];
      if (sub < a.planes) load(nxt, sub);
      for (int p = sub; p < a.planes; p += S) {
        const float f = a.mom.f[p], f2 = a.mom.f2[p];
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = nxt[m];
        if (p + S < a.planes) load(nxt, p + S);
        F::template run<false>(v, ws, twl, lane);
        acc.add(v, f, f2, p < a.mom.split, a.mom.power2);
      }
    } else if (active) {  // more waves per CU instead of a register prefetch
      for (int p = sub; p < a.planes; p += S) {
        const float f = a.mom.f[p], f2 = a.mom.f2[p];
        float2 v[E];
        load(v, p);
        F::template run<false>(v, ws, twl, lane);
        acc.add(v, f, f2, p < a.mom.split, a.mom.power2);
      }
    }
    // fixed-order combination of the plane subsets: sub 0 += sub 1, sub 2, ... (LDS reused)
    float *red = reinterpret_cast<float *>(scr);  // [(S-1)][4 columns][N]
    __syncthreads();
    // k is a compile-time index after unrolling: a runtime index into acc.s would move
    // the sums to scratch memory
#pragma unroll
    for (int k = 0; k < NACC; ++k) {
      if (S == 1) continue;
      if (sub > 0 && active) {
#pragma unroll
        for (int m = 0; m < E; ++m) red[((sub - 1) * 4 + c) * N + lane + 64 * m] = acc.s[k][m];
      }
      __syncthreads();
      if (sub == 0 && active) {
        for (int s = 1; s < S; ++s)
#pragma unroll
          for (int m = 0; m < E; ++m) acc.s[k][m] += red[((s - 1) * 4 + c) * N + lane + 64 * m];
      }
      __syncthreads();
    }
    if (sub == 0 && active) {
#pragma unroll
      for (int m = 0; m < E; ++m) {
        float val[5] = {acc.s[0][m], acc.s[1][m], acc.s[2][m], 0.f, 0.f};
        if constexpr (NACC == 5) {
          val[3] = acc.s[3][m];
          val[4] = acc.s[4][m];
        }
        store_moment(a, pw.mom_out, NACC, lane + 64 * m, x, val);
      }
    }
  } else {
    if (!active) return;
    // one plane's column: transform (BLK_MID: forward, times the table, inverse) and store
    auto finish = [&](float2 (&v)[E], long long p) {
      if (KIND == BLK_MID) {
        F::template run<false>(v, ws, twl, lane);
        asm volatile("" ::: "memory");  // keep the table loads after the forward transform
#pragma unroll
        for (int m = 0; m < E; ++m) v[m] = cmul(v[m], pw.mid_tf[tile + lane * C + c + m * col_step]);
        F::template run<true>(v, ws, twl, lane);
#pragma unroll
        for (int m = 0; m < E; ++m)
          reinterpret_cast<float2 *>(pw.out)[p * plane_elems + tile + lane * C + c + m * col_step] = v[m];
      } else {
        F::template run<false>(v, ws, twl, lane);
#pragma unroll
        for (int m = 0; m < E; ++m) {
          const int y = lane + 64 * m;
          float2 t = v[m];
          const long long o = p * plane_elems + (long long)y * a.W + x;
          if (a.out_mode == OUT_NAT_MAG) {
            reinterpret_cast<float *>(pw.out)[o] = sqrtf(t.x * t.x + t.y * t.y);
          } else if (a.out_mode == OUT_NAT_RE) {
            reinterpret_cast<float *>(pw.out)[o] = t.x;
          } else {
            if (a.post_lx) t = cmul(t, cmul(a.post_lx[x], a.post_ly[y]));
            reinterpret_cast<float2 *>(pw.out)[o] = t;
          }
        }
      }
    };
    float2 v[E];
    load(v, bp);
    finish(v, bp);
  }
}

// ------------------------------------------------------------------------- launch
#define HFX_WAVE_SIZES(X) X(256) X(512) X(768) X(1024) X(2048)

bool wave_pass_supported(int n) { return n == 256 || n == 512 || n == 768 || n == 1024 || n == 2048; }

template <int N, int KIND> static int launch_wrow_n(const PassArgs &a, hipStream_t st) {
  if ((a.inv != 0) != row_inv(KIND)) return -5;
  constexpr int RW = wrow_waves(N, KIND);
  const unsigned gx = (unsigned)((a.H + RW - 1) / RW);
  const unsigned gy = is_mom(KIND) ? 1u : (unsigned)a.planes;
  const bool plain = KIND == NAT_TO_BLK && a.in_mode == IN_NAT_C && a.in_h == a.H && a.in_w == a.W &&
                     a.off_x == 0 && a.off_y == 0 && a.pre_lx == nullptr;
  if (a.nwin > 1)
    hipLaunchKernelGGL((k_wrow<N, KIND, true>), dim3(gx, gy, (unsigned)a.nwin), dim3(64 * RW), 0, st, a);
  else if (plain)
    hipLaunchKernelGGL((k_wrow<N, KIND, false, KIND == NAT_TO_BLK>), dim3(gx, gy), dim3(64 * RW), 0, st, a);
  else
    hipLaunchKernelGGL((k_wrow<N, KIND>), dim3(gx, gy), dim3(64 * RW), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int N, int KIND> static int launch_wcol_n(const PassArgs &a, hipStream_t st) {
  if (a.inv != 0) return -5;
  constexpr int TL = wcol_tiles(N, KIND);
  const unsigned tiles = (unsigned)((a.W / a.blk_c + TL - 1) / TL);
  // BLK_MID: one flat dimension, plane-fastest (see k_wcol); the plane-parallel others: y = plane
  const unsigned gx = KIND == BLK_MID ? tiles * (unsigned)a.planes : tiles;
  const unsigned gy = is_mom(KIND) || KIND == BLK_MID ? 1u : (unsigned)a.planes;
  const unsigned nt = wcol_threads(N, KIND);
  if (a.nwin > 1)
    hipLaunchKernelGGL((k_wcol<N, KIND, true>), dim3(gx, gy, (unsigned)a.nwin), dim3(nt), 0, st, a);
  else
    hipLaunchKernelGGL((k_wcol<N, KIND>), dim3(gx, gy), dim3(nt), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <int KIND> static int wrow_dispatch(const PassArgs &a, hipStream_t st) {
#define X(s) if (a.W == s) return launch_wrow_n<s, KIND>(a, st);
  HFX_WAVE_SIZES(X)
#undef X
  return -2;
}

template <int KIND> static int wcol_dispatch(const PassArgs &a, hipStream_t st) {
#define X(s) if (a.H == s) return launch_wcol_n<s, KIND>(a, st);
  HFX_WAVE_SIZES(X)
#undef X
  return -2;
}

int launch_wave_row_pass(const PassArgs &a, hipStream_t st) {
  if (a.in_mode != IN_BLK) {
    if (a.out_mode != OUT_BLK) return -4;
    return wrow_dispatch<NAT_TO_BLK>(a, st);
  }
  if (a.out_mode == OUT_MOMENTS)
    return a.mom.split > 0 ? wrow_dispatch<BLK_TO_MOM5>(a, st) : wrow_dispatch<BLK_TO_MOM>(a, st);
  if (a.out_mode == OUT_NAT_C || a.out_mode == OUT_NAT_MAG || a.out_mode == OUT_NAT_RE) return wrow_dispatch<BLK_TO_NAT>(a, st);
  return -4;
}

int launch_wave_col_pass(const PassArgs &a, hipStream_t st) {
  if (a.in_mode != IN_BLK || a.blk_c < 1 || a.blk_c > 4 || a.W % a.blk_c != 0) return -3;
  if (a.mid_tf) {
    if (a.out_mode != OUT_BLK) return -4;
    return wcol_dispatch<BLK_MID>(a, st);
  }
  if (a.out_mode == OUT_MOMENTS)
    return a.mom.split > 0 ? wcol_dispatch<BLK_TO_MOM5>(a, st) : wcol_dispatch<BLK_TO_MOM>(a, st);
  if (a.out_mode == OUT_NAT_C || a.out_mode == OUT_NAT_MAG || a.out_mode == OUT_NAT_RE) return wcol_dispatch<BLK_TO_NAT>(a, st);
  return -4;
}

}  // namespace hfx
