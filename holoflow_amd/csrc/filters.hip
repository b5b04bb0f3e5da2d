// Optional in-path stages: the Filter2D band-pass table, the Flatfield Gaussian high-pass
// (Fourier table and the spatial separable form).
//
// Both spectral stages run on the shared 2-D FFT passes (row pass, column pass with the table
// multiplied between its forward and inverse transforms, inverse row pass); these kernels
// build their tables, in the column-blocked layout those passes read:
//   (r, c) -> (c / blk_c) * H * blk_c + r * blk_c + c % blk_c.
#include "hfx_kernels.hh"

namespace hfx {

namespace {

constexpr float kPiF = 3.14159265358979f;

__device__ __forceinline__ long long tab_index(int r, int c, int H, int W, int blk_c) {
  return blk_c > 0 ? (long long)(c / blk_c) * H * blk_c + (long long)r * blk_c + (c % blk_c)
                   : (long long)r * W + c;
}

// Cosine-tapered edge of the annulus (filter2d.cu:165-240): 1 inside radius r, a quarter
// cosine over the next s pixels, 0 beyond.  `q` is the squared distance scaled by `k2`
// (4 for the half-integer centred grid, 1 for the integer grid), so that the comparisons
// against r and r + s are exact integer tests; the taper itself is evaluated in FP32.
__device__ __forceinline__ float taper(long long q, int k2, float dist, int r, int s) {
  if (q < (long long)k2 * r * r) return 1.0f;
  const long long rs = (long long)r + s;
  if (s > 0 && q < (long long)k2 * rs * rs)
    return cosf((dist - (float)r) / (float)s * kPiF / 2);
  return 0.0f;
}

}  // namespace

// Filter2D table (filter2d.cu:165-305).
//   complex input: the annulus a(1-b) around the grid centre (W/2.0, H/2.0), stored after the
//     quadrant exchange of the four corners (DC to index 0 for even sizes), with the value
//     in both the real and the imaginary part, as the reference's table holds it;
//   real input: the annulus around DC on the integer frequency grid (|kx|, signed ky), real.
// Every table entry is the one the reference's kernels write for that frequency; the real
// input's half spectrum is extended to the full grid through |kx|.
__global__ void k_filter2d_table(float2 *out, int H, int W, int r_in, int r_out, int s_in, int s_out,
                                 int complex_mode, int blk_c) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= W || r >= H) return;
  long long q;
  int k2;
  if (complex_mode) {
    // position before the corner exchange (only the 2*(W/2) x 2*(H/2) block moves)
    const int hw = W / 2, hh = H / 2;
    int rs = r, cs = c;
    if (r < 2 * hh && c < 2 * hw) {
      rs = r < hh ? r + hh : r - hh;
      cs = c < hw ? c + hw : c - hw;
    }
    const long long x2 = 2LL * cs - W, y2 = 2LL * rs - H;  // twice the centred coordinates
    q = x2 * x2 + y2 * y2;
    k2 = 4;
  } else {
    const long long kx = c <= W / 2 ? c : W - c;
    const long long ky = r <= H / 2 ? r : r - H;
    q = kx * kx + ky * ky;
    k2 = 1;
  }
  const float dist = (float)(sqrt((double)q) / (k2 == 4 ? 2.0 : 1.0));
  const float a = taper(q, k2, dist, r_out, s_out);
  const float b = taper(q, k2, dist, r_in, s_in);
  const float m = a * (1 - b);
  out[tab_index(r, c, H, W, blk_c)] = make_float2(m, complex_mode ? m : 0.0f);
}

void launch_filter2d_table(float2 *out, int H, int W, int r_in, int r_out, int s_in, int s_out,
                           int complex_mode, int blk_c, hipStream_t st) {
  hipLaunchKernelGGL(k_filter2d_table, dim3((W + 127) / 128, H), dim3(128), 0, st, out, H, W, r_in, r_out,
                     s_in, s_out, complex_mode, blk_c);
}

// Flatfield Gaussian high-pass in the Fourier domain (flatfield.cu:143-180):
//   gain(kx, ky) = (1 - exp(-2 pi^2 (sx^2 fx^2 + sy^2 fy^2))) * scale,
//   fx = |kx| / W, fy = signed ky / H, real; evaluated in FP32 as the reference does.
__global__ void k_flatfield_table(float2 *out, int H, int W, float sigma_y, float sigma_x, float scale,
                                  int blk_c) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (c >= W || r >= H) return;
  const int kx = c <= W / 2 ? c : W - c;
  const float fx = (float)kx / (float)W;
  const float fy_idx = r <= H / 2 ? (float)r : (float)(r - H);
  const float fy = fy_idx / (float)H;
  const float two_pi_squared = 19.739208802178716f;
  const float lowpass = expf(-two_pi_squared * (sigma_x * sigma_x * fx * fx + sigma_y * sigma_y * fy * fy));
  out[tab_index(r, c, H, W, blk_c)] = make_float2((1.0f - lowpass) * scale, 0.0f);
}

void launch_flatfield_table(float2 *out, int H, int W, float sigma_y, float sigma_x, float scale, int blk_c,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_flatfield_table, dim3((W + 127) / 128, H), dim3(128), 0, st, out, H, W, sigma_y,
                     sigma_x, scale, blk_c);
}

// Flatfield, spatial form (flatfield.cu:189-240): background = Gaussian blur with clamp-at-
// edge indexing, horizontal along x then vertical along y, out = in - background; the taps
// are summed in ascending order with fused multiply-adds.
//
// One block owns an FF_TY x FF_TX tile of one plane: it stages the tile and its clamped
// halo in LDS, blurs the halo rows horizontally into LDS, then blurs vertically and
// subtracts, so each input element is read from HBM about once (halo re-reads hit L2).
constexpr int FF_TX = 64, FF_TY = 16, FF_RMAX = 31, FF_NT = 256;

// The tile buffers are sized for the actual radii (dynamic LDS, ff_spatial_lds_bytes): at
// sigma 3 (radius 9) a block holds 20 KB instead of the 59 KB of the largest radius, so
// eight blocks share a CU instead of two.
__host__ __device__ constexpr size_t ff_spatial_lds_bytes(int rx, int ry) {
  return (size_t)(FF_TY + 2 * ry) * ((FF_TX + 2 * rx) + FF_TX) * sizeof(float);
}

__global__ __launch_bounds__(FF_NT) void k_flatfield_spatial(const float *__restrict__ in, float *__restrict__ out,
                                                              int H, int W, const float *__restrict__ kx, int rx,
                                                              const float *__restrict__ ky, int ry) {
  extern __shared__ float ff_lds[];
  __shared__ float wx[2 * FF_RMAX + 1], wy[2 * FF_RMAX + 1];
  const int tid = threadIdx.x;
  const int x0 = blockIdx.x * FF_TX, y0 = blockIdx.y * FF_TY;
  const long long base = (long long)blockIdx.z * H * W;
  const int rows = FF_TY + 2 * ry, pitch = FF_TX + 2 * rx;
  float *raw = ff_lds;                 // [rows][pitch]
  float *hor = ff_lds + rows * pitch;  // [rows][FF_TX]
  for (int i = tid; i <= 2 * rx; i += FF_NT) wx[i] = kx[i];
  for (int i = tid; i <= 2 * ry; i += FF_NT) wy[i] = ky[i];
  for (int i = tid; i < rows * pitch; i += FF_NT) {
    const int rr = i / pitch, cc = i - rr * pitch;
    const int y = min(max(y0 - ry + rr, 0), H - 1);
    const int x = min(max(x0 - rx + cc, 0), W - 1);
    raw[i] = in[base + (long long)y * W + x];
  }
  __syncthreads();
  for (int i = tid; i < rows * FF_TX; i += FF_NT) {
    const int rr = i / FF_TX, cc = i - rr * FF_TX;
    const float *row = raw + rr * pitch + cc;
    float s = 0.0f;
    for (int k = 0; k <= 2 * rx; ++k) s = fmaf(row[k], wx[k], s);
    hor[i] = s;
  }
  __syncthreads();
  for (int i = tid; i < FF_TY * FF_TX; i += FF_NT) {
    const int rr = i / FF_TX, cc = i - rr * FF_TX;
    const int y = y0 + rr, x = x0 + cc;
    if (y >= H || x >= W) continue;
    float b = 0.0f;
    for (int k = 0; k <= 2 * ry; ++k) b = fmaf(hor[(rr + k) * FF_TX + cc], wy[k], b);
    out[base + (long long)y * W + x] = raw[(rr + ry) * pitch + cc + rx] - b;
  }
}

// Radii beyond the LDS tile's halo (only on one-pixel-wide or -tall planes, where the
// Fourier form is not used): the same sums as two global passes through `tmp`.
__global__ void k_flatfield_rows(const float *__restrict__ in, float *__restrict__ tmp, long long total, int H,
                                 int W, const float *__restrict__ kx, int rx) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long long plane = (long long)H * W;
  const long long base = idx / plane * plane;
  const int local = (int)(idx - base), y = local / W, x = local - y * W;
  float s = 0.0f;
  for (int k = -rx; k <= rx; ++k) s = fmaf(in[base + (long long)y * W + min(max(x + k, 0), W - 1)], kx[k + rx], s);
  tmp[idx] = s;
}

__global__ void k_flatfield_cols(const float *__restrict__ in, const float *__restrict__ tmp,
                                 float *__restrict__ out, long long total, int H, int W,
                                 const float *__restrict__ ky, int ry) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long long plane = (long long)H * W;
  const long long base = idx / plane * plane;
  const int local = (int)(idx - base), y = local / W, x = local - y * W;
  float b = 0.0f;
  for (int k = -ry; k <= ry; ++k) b = fmaf(tmp[base + (long long)min(max(y + k, 0), H - 1) * W + x], ky[k + ry], b);
  out[idx] = in[idx] - b;
}

bool flatfield_spatial_fused(int rx, int ry) { return rx <= FF_RMAX && ry <= FF_RMAX; }

int launch_flatfield_spatial(const float *in, float *out, float *tmp, long long planes, int H, int W,
                             const float *kx, int rx, const float *ky, int ry, hipStream_t st) {
  if (flatfield_spatial_fused(rx, ry)) {
    if (planes > 65535) return -1;
    hipLaunchKernelGGL(k_flatfield_spatial, dim3((W + FF_TX - 1) / FF_TX, (H + FF_TY - 1) / FF_TY, (unsigned)planes),
                       dim3(FF_NT), ff_spatial_lds_bytes(rx, ry), st, in, out, H, W, kx, rx, ky, ry);
  } else {
    if (!tmp) return -1;
    const long long total = planes * H * W;
    const unsigned grid = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(k_flatfield_rows, dim3(grid), dim3(256), 0, st, in, tmp, total, H, W, kx, rx);
    hipLaunchKernelGGL(k_flatfield_cols, dim3(grid), dim3(256), 0, st, in, tmp, out, total, H, W, ky, ry);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hfx
