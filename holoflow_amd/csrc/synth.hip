// Seeded synthetic laser-Doppler interferograms, generated on the GPU.
//
// Random numbers: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), a counter-based
// generator, keyed by (seed, stream_id) and counted by (pixel, frame, purpose, 0), so any
// frame of any pixel can be produced independently and the CPU twin in the oracle
// produces the same draws.  Normals come from Box-Muller on two 32-bit words.
//
// Model, per pixel p = (y, x), frame n (all arithmetic in FP64):
//   object   O(p)  = obj_sigma * (g0 + i g1)/sqrt(2) inside the centred disk, else 0
//   Doppler  fd(p) = fd_max * (0.5 + 0.5 sin(2 pi (1.3 x/W + 0.17)) cos(2 pi (0.9 y/H - 0.05)))
//   phase    phi_n = wrap(phi_{n-1} + 2 pi fd/fs + phase_noise * g_n),  phi_{-1} = 0
//   ref beam R(p)  = ref_amp * exp(2 pi i carrier x)
//   counts   I     = gain |R + O e^{i phi_n}|^2,  I' = I + sqrt(I) * h_n   (shot-like noise)
//   frame    u16   = clamp(rint(I'), 0, 4095)
#include "hfx_kernels.hh"

namespace hfx {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo0 = 0xD2511F53u * c[0];
    uint32_t hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t lo1 = 0xCD9E8D57u * c[2];
    uint32_t n0 = hi1 ^ c[1] ^ k0;
    uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ double u01(uint32_t u) { return ((double)u + 0.5) * (1.0 / 4294967296.0); }

__device__ __forceinline__ void box_muller(uint32_t a, uint32_t b, double &g0, double &g1) {
  double r = sqrt(-2.0 * log(u01(a)));
  double s, c;
  sincospi(2.0 * u01(b), &s, &c);
  g0 = r * c;
  g1 = r * s;
}

__global__ void k_synth(SynthArgs a, uint16_t *frames, long long first, long long nf,
                        double *phase_state) {
  long long npix = (long long)a.H * a.W;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npix) return;
  int y = (int)(p / a.W), x = (int)(p % a.W);
  double cx = 0.5 * (a.W - 1), cy = 0.5 * (a.H - 1);
  bool inside = (x - cx) * (x - cx) + (y - cy) * (y - cy) < a.disk_radius * a.disk_radius;
  double ore = 0.0, oim = 0.0;
  if (inside) {
    uint32_t c[4] = {(uint32_t)p, 0u, 0u, 0xA0u};
    philox4x32_10(c, a.seed, a.stream_id);
    double g0, g1;
    box_muller(c[0], c[1], g0, g1);
    ore = a.obj_sigma * g0 * 0.70710678118654752440;
    oim = a.obj_sigma * g1 * 0.70710678118654752440;
  }
  double fd = a.fd_max * (0.5 + 0.5 * sinpi(2.0 * (1.3 * x / a.W + 0.17)) *
                                    cospi(2.0 * (0.9 * y / a.H - 0.05)));
  double dphi = 2.0 * M_PI * fd / a.fs;
  double rs, rc;
  sincospi(2.0 * a.carrier * x, &rs, &rc);
  double rre = a.ref_amp * rc, rim = a.ref_amp * rs;
  double phi = phase_state[p];
  for (long long i = 0; i < nf; ++i) {
    long long n = first + i;
    uint32_t c[4] = {(uint32_t)p, (uint32_t)n, 1u, 0xB0u};
    philox4x32_10(c, a.seed, a.stream_id);
    double g, h;
    box_muller(c[0], c[1], g, h);
    phi = phi + dphi + a.phase_noise * g;
    phi = phi - 2.0 * M_PI * floor((phi + M_PI) / (2.0 * M_PI));  // wrap to [-pi, pi)
    double s, co;
    sincos(phi, &s, &co);
    double fre = rre + ore * co - oim * s;
    double fim = rim + ore * s + oim * co;
    double I = a.gain * (fre * fre + fim * fim);
    double v = rint(I + sqrt(I) * h);
    v = v < 0.0 ? 0.0 : (v > 4095.0 ? 4095.0 : v);
    frames[i * npix + p] = (uint16_t)v;
  }
  phase_state[p] = phi;
}

int launch_synth(const SynthArgs &a, uint16_t *frames, long long first_frame, long long n_frames,
                 double *phase_state, hipStream_t st) {
  long long npix = (long long)a.H * a.W;
  hipLaunchKernelGGL(k_synth, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, st, a, frames,
                     first_frame, n_frames, phase_state);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace hfx
