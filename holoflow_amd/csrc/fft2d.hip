// 2-D FFT passes for the propagation stage (Fresnel and angular spectrum), each an
// LDS-staged Stockham transform along one axis with its neighbours fused in.
//
//   row pass    : tile = ROWS consecutive rows of length N (= W)
//   column pass : tile = CL LDS columns of length N (= H), fed from the column-blocked
//                 layout so that every global access is a contiguous run
//
// Kernel kinds (compile-time, so each kernel carries only its own data path):
//   NAT_TO_BLK : natural complex or real input (optionally centred in a zero-padded grid
//                and times the Fresnel lens) -> column-blocked spectrum      (row pass)
//   BLK_TO_MOM : column-blocked input -> transform -> moment images, looping over all
//                planes of the tile (row pass inverse for the angular spectrum, column
//                pass forward for Fresnel)
//   BLK_MID    : column-blocked -> forward, times the transfer function, inverse ->
//                column-blocked in place                                      (column pass)
//   BLK_TO_NAT : column-blocked -> transform -> natural complex (optionally times the
//                output lens) or natural |X|
//
// Column-blocked layout of one H x W plane with block width C (C | W):
//   (y, x) -> (x / C) * H * C + y * C + x % C
// A row pass writes ROWS*C contiguous elements per block; a column pass reads H*C.
//
// Moment epilogue (graph_builder.cc:604-647, mean_abs.cu:239-276, mean.cu:268-306):
//   a_k = |X_k| (or |X_k|^2), M0 = sum_k a_k / n, M1 = sum_k (f_k a_k) / n,
//   M2 = sum_k (f_k^2 a_k) / n, summed in ascending k with the products rounded before
//   the sum as the reference's multiply-then-mean does; stored with the optional
//   fftshift out[i] = in[(i + n/2) mod n] (fftshift.cu:110-145).
#include "hfx_kernels.hh"

namespace hfx {

enum Kind { NAT_TO_BLK = 0, BLK_TO_MOM = 1, BLK_MID = 2, BLK_TO_NAT = 3 };

__device__ __forceinline__ long long blk_index(int y, int x, int H, int C) {
  return (long long)(x / C) * H * C + (long long)y * C + (x % C);
}

constexpr int row_tile_rows(int n) {
  int r = 4096 / n;
  return r < 1 ? 1 : (r > 64 ? 64 : r);
}
constexpr int pass_threads(int tile) {
  int t = tile / 8;
  t = (t + 63) / 64 * 64;
  return t < 64 ? 64 : (t > 1024 ? 1024 : t);
}
// LDS columns of a column-pass tile.  Four columns give W/4 tiles per plane (256 for a
// 1024-wide plane, one per CU) while each tile is still one contiguous run of H*4*8 bytes.
constexpr int col_lds_width(int) { return 4; }

template <int N, bool COLS> struct Geo {
  static constexpr int CL = col_lds_width(N);
  static constexpr int ROWS = COLS ? CL : row_tile_rows(N);
  static constexpr int TILE = N * ROWS;
  static constexpr int NT = pass_threads(TILE);
  static constexpr int E = (TILE + NT - 1) / NT;
  static constexpr int RS = COLS ? 1 : N;
  static constexpr int ES = COLS ? CL : 1;
};

// Element e of a tile -> LDS slot li and grid coordinates (y, x).
// blk_order selects the enumeration whose consecutive e are contiguous in the blocked
// layout; otherwise consecutive e are contiguous in the natural layout.
template <int N, bool COLS>
__device__ __forceinline__ void tile_elem(int e, bool blk_order, int C, int y0, int x0, int &li,
                                          int &y, int &x) {
  using G = Geo<N, COLS>;
  if (COLS) {
    int i = e / G::CL, c = e % G::CL;
    li = i * G::CL + c;
    y = i;
    x = (c < C) ? x0 + c : -1;  // LDS columns beyond the block width are padding
  } else if (blk_order) {
    int c = e % C;
    int row = (e / C) % G::ROWS;
    int blk = e / (C * G::ROWS);
    li = row * N + blk * C + c;
    y = y0 + row;
    x = blk * C + c;
  } else {
    li = e;
    y = y0 + e / N;
    x = e % N;
  }
}

template <int N, bool COLS, int KIND>
__global__ __launch_bounds__((Geo<N, COLS>::NT)) void k_pass(PassArgs a) {
  using G = Geo<N, COLS>;
  using Fft = LdsFft<N, G::ROWS, G::RS, G::ES, G::NT>;
  constexpr bool MOM = KIND == BLK_TO_MOM;
  __shared__ float2 lds[G::TILE];
  __shared__ float2 twl[N];
  const int tid = threadIdx.x;
  for (int i = tid; i < N; i += G::NT) twl[i] = a.tw[i];

  const int C = a.blk_c;
  const int y0 = COLS ? 0 : xcd_tile(blockIdx.x, gridDim.x) * G::ROWS;
  const int x0 = COLS ? blockIdx.x * C : 0;
  const int P = MOM ? a.planes : 1;
  const long long pbase = MOM ? 0 : blockIdx.y;
  const long long plane_elems = (long long)a.H * a.W;
  const bool split = MOM && a.mom.split > 0;

  float acc[MOM ? 5 : 1][MOM ? G::E : 1];
  if constexpr (MOM) {
#pragma unroll
    for (int k = 0; k < 5; ++k)
#pragma unroll
      for (int it = 0; it < G::E; ++it) acc[k][it] = 0.f;
  }

  // Global loads of one plane's tile into registers; the next plane's loads are issued
  // before the current plane's transform so their latency hides behind the LDS work.
  float2 pre[G::E];
  auto load_plane = [&](long long p, int t) {
#pragma unroll
    for (int it = 0; it < G::E; ++it) {
      int e = t + it * G::NT;
      pre[it] = make_float2(0.f, 0.f);
      if (e < G::TILE) {
        int li, y, x;
        tile_elem<N, COLS>(e, KIND != NAT_TO_BLK, C, y0, x0, li, y, x);
        if (y < a.H && x >= 0) {
          if (KIND == NAT_TO_BLK) {
            const int yi = y - a.off_y, xi = x - a.off_x;
            if (yi >= 0 && yi < a.in_h && xi >= 0 && xi < a.in_w) {
              const long long o = p * (long long)a.in_h * a.in_w + (long long)yi * a.in_w + xi;
              float2 v;
              if (a.in_mode == IN_NAT_C)
                v = reinterpret_cast<const float2 *>(a.in)[o];
              else
                v = make_float2(reinterpret_cast<const float *>(a.in)[o], 0.f);
              if (a.pre_lx) v = cmul(v, cmul(a.pre_lx[xi], a.pre_ly[yi]));
              pre[it] = v;
            }
          } else {
            pre[it] = reinterpret_cast<const float2 *>(a.in)[p * plane_elems + blk_index(y, x, a.H, C)];
          }
        }
      }
    }
  };
  load_plane(pbase, tid);

  for (int pp = 0; pp < P; ++pp) {
    const long long p = pbase + pp;
    // An opaque copy of the thread index per plane keeps the per-element index arithmetic
    // inside the loop instead of being hoisted into registers for all planes.
    int t;
    asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"(tid));

    // ---------------------------------------------------------------- registers -> LDS
#pragma unroll
    for (int it = 0; it < G::E; ++it) {
      int e = t + it * G::NT;
      if (e < G::TILE) {
        int li, y, x;
        tile_elem<N, COLS>(e, KIND != NAT_TO_BLK, C, y0, x0, li, y, x);
        lds[li] = pre[it];
      }
    }
    __syncthreads();
    if (pp + 1 < P) load_plane(p + 1, t);

    // ---------------------------------------------------------------- transform(s)
    if (KIND == BLK_MID) {
      Fft::run(lds, twl, t, false);
#pragma unroll
      for (int it = 0; it < G::E; ++it) {
        int e = t + it * G::NT;
        if (e < G::TILE) {
          int li, y, x;
          tile_elem<N, COLS>(e, true, C, y0, x0, li, y, x);
          if (y < a.H && x >= 0) lds[li] = cmul(lds[li], a.mid_tf[blk_index(y, x, a.H, C)]);
        }
      }
      __syncthreads();
      Fft::run(lds, twl, t, true);
    } else {
      Fft::run(lds, twl, t, a.inv != 0);
    }

    // ---------------------------------------------------------------- store / accumulate
    const float f = MOM ? a.mom.f[pp] : 0.f;
    const float f2 = MOM ? a.mom.f2[pp] : 0.f;
#pragma unroll
    for (int it = 0; it < G::E; ++it) {
      int e = t + it * G::NT;
      if (e < G::TILE) {
        int li, y, x;
        tile_elem<N, COLS>(e, KIND == NAT_TO_BLK || KIND == BLK_MID, C, y0, x0, li, y, x);
        if (y < a.H && x >= 0) {
          float2 v = lds[li];
          if constexpr (MOM) {
            float m = sq_mag(v);
            if (!a.mom.power2) m = mag_sqrt(m);
            acc[0][it] += m;
            acc[1][it] += __fmul_rn(f, m);
            acc[2][it] += __fmul_rn(f2, m);
            if (split) {
              if (pp < a.mom.split)
                acc[3][it] += m;
              else
                acc[4][it] += m;
            }
          } else if (KIND == NAT_TO_BLK || KIND == BLK_MID) {
            reinterpret_cast<float2 *>(a.out)[p * plane_elems + blk_index(y, x, a.H, C)] = v;
          } else {
            const long long o = p * plane_elems + (long long)y * a.W + x;
            if (a.out_mode == OUT_NAT_MAG) {
              reinterpret_cast<float *>(a.out)[o] = sqrtf(v.x * v.x + v.y * v.y);
            } else if (a.out_mode == OUT_NAT_RE) {
              reinterpret_cast<float *>(a.out)[o] = v.x;
            } else {
              if (a.post_lx) v = cmul(v, cmul(a.post_lx[x], a.post_ly[y]));
              reinterpret_cast<float2 *>(a.out)[o] = v;
            }
          }
        }
      }
    }
    if (pp + 1 < P) __syncthreads();
  }

  if constexpr (MOM) {
    const float n_all = (float)P;
    const float n_lo = (float)a.mom.split, n_hi = (float)(P - a.mom.split);
#pragma unroll
    for (int it = 0; it < G::E; ++it) {
      int e = tid + it * G::NT;
      if (e < G::TILE) {
        int li, y, x;
        tile_elem<N, COLS>(e, false, C, y0, x0, li, y, x);
        if (y < a.H && x >= 0) {
          int ys = y, xs = x;
          if (a.mom.shift) {
            ys = (y + a.H - a.H / 2) % a.H;
            xs = (x + a.W - a.W / 2) % a.W;
          }
          const long long o = (long long)ys * a.W + xs;
          a.mom.out[0][o] = acc[0][it] / n_all;
          a.mom.out[1][o] = acc[1][it] / n_all;
          a.mom.out[2][o] = acc[2][it] / n_all;
          if (split) {
            if (a.mom.out[3]) a.mom.out[3][o] = acc[3][it] / n_lo;
            if (a.mom.out[4]) a.mom.out[4][o] = acc[4][it] / n_hi;
          }
        }
      }
    }
  }
}

#define HFX_SIZES(X) \
  X(2) X(4) X(8) X(16) X(24) X(32) X(48) X(64) X(96) X(128) X(256) X(512) X(768) X(1024) X(2048)

bool row_pass_supported(int W) {
#define X(n) if (W == n) return true;
  HFX_SIZES(X)
#undef X
  return false;
}

int pick_blk_c(int W, int H) {
  for (int c = col_lds_width(H); c > 1; c /= 2)
    if (W % c == 0) return c;
  return 1;
}

bool col_pass_supported(int H, int blk_c) {
  return row_pass_supported(H) && blk_c >= 1 && blk_c <= col_lds_width(H);
}

template <int N, bool COLS, int KIND> static int launch_n(const PassArgs &a, hipStream_t st) {
  using G = Geo<N, COLS>;
  unsigned gx = COLS ? (unsigned)(a.W / a.blk_c) : (unsigned)((a.H + G::ROWS - 1) / G::ROWS);
  unsigned gy = KIND == BLK_TO_MOM ? 1u : (unsigned)a.planes;
  hipLaunchKernelGGL((k_pass<N, COLS, KIND>), dim3(gx, gy), dim3(G::NT), 0, st, a);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

template <bool COLS, int KIND> static int dispatch(const PassArgs &a, int n, hipStream_t st) {
#define X(s) if (n == s) return launch_n<s, COLS, KIND>(a, st);
  HFX_SIZES(X)
#undef X
  return -2;
}

int launch_row_pass(const PassArgs &a, hipStream_t st) {
  if (wave_pass_supported(a.W)) return launch_wave_row_pass(a, st);
  if (a.in_mode != IN_BLK) {
    if (a.out_mode != OUT_BLK) return -4;
    return dispatch<false, NAT_TO_BLK>(a, a.W, st);
  }
  if (a.out_mode == OUT_MOMENTS) return dispatch<false, BLK_TO_MOM>(a, a.W, st);
  if (a.out_mode == OUT_NAT_C || a.out_mode == OUT_NAT_MAG || a.out_mode == OUT_NAT_RE) return dispatch<false, BLK_TO_NAT>(a, a.W, st);
  return -4;
}

int launch_col_pass(const PassArgs &a, hipStream_t st) {
  if (a.in_mode != IN_BLK || a.W % a.blk_c != 0 || !col_pass_supported(a.H, a.blk_c)) return -3;
  if (wave_pass_supported(a.H)) return launch_wave_col_pass(a, st);
  if (a.mid_tf) {
    if (a.out_mode != OUT_BLK) return -4;
    return dispatch<true, BLK_MID>(a, a.H, st);
  }
  if (a.out_mode == OUT_MOMENTS) return dispatch<true, BLK_TO_MOM>(a, a.H, st);
  if (a.out_mode == OUT_NAT_C || a.out_mode == OUT_NAT_MAG || a.out_mode == OUT_NAT_RE) return dispatch<true, BLK_TO_NAT>(a, a.H, st);
  return -4;
}

}  // namespace hfx
