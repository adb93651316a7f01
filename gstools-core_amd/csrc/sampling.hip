// On-device seeded sampling of spectral modes (the caller side of the
// summators: GSTools' RandMeth draws the wave vectors from the model's
// spectral density and z1, z2 ~ N(0, 1), gen_benchmark_inputs.py:17-29).
//
// Random bits: Philox4x64-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC 2011), the counter-based generator of
// numpy.random.Philox, with the same key and counter conventions: key =
// (seed, stream), and the b-th block of four 64-bit words (b = 0, 1, ...)
// is the cipher of the counter (b + 1, 0, 0, 0).  Word i of the stream is
// therefore word i of numpy.random.Philox with key words (seed, stream).
//
// Mode m of a realisation uses blocks 2m and 2m + 1 (eight words), so a mode
// depends only on (seed, stream, m), never on M or on the thread that draws
// it.  Word w gives the uniform u = (w >> 11) 2^-53 in [0, 1); the pairs
// (u0, u1), (u2, u3), (u4, u5), (u6, u7) give eight standard normals by
// Box-Muller, n = sqrt(-2 ln(1 - u_even)) (cos, sin)(2 pi u_odd), all FP64:
//   g_d = n_d (d < dim), z1 = n_4, z2 = n_5, w = n_6.
// Wave vectors:
//   Gaussian model:     k = g / l                (k ~ N(0, l^-2 I): the spectral
//                       law of the covariance exp(-r^2 / (2 l^2)))
//   exponential model:  k = g / (l |w|)          (isotropic multivariate t, 1 dof:
//                       the spectral law of exp(-r / l); |w| is kept >= DBL_MIN,
//                       so k stays finite)
#include "gsc_device.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

namespace gsc {
namespace {

constexpr uint64_t kPhiloxM0 = 0xD2E7470EE14C6C93ull;
constexpr uint64_t kPhiloxM1 = 0xCA5A826395121157ull;
constexpr uint64_t kPhiloxW0 = 0x9E3779B97F4A7C15ull;  // golden ratio
constexpr uint64_t kPhiloxW1 = 0xBB67AE8584CAA73Bull;  // sqrt(3) - 1

__device__ __forceinline__ void philox4x64_10(uint64_t (&c)[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t hi0 = __umul64hi(kPhiloxM0, c[0]), lo0 = kPhiloxM0 * c[0];
    const uint64_t hi1 = __umul64hi(kPhiloxM1, c[2]), lo1 = kPhiloxM1 * c[2];
    const uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += kPhiloxW0;
    k1 += kPhiloxW1;
  }
}

__global__ void __launch_bounds__(256) k_sample_modes(SampleArgs a) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (t >= a.m_total) return;
  int r = 0;
  int64_t m0 = 0, M = a.m_total;
  double ls = a.len_scale1;
  uint64_t seed = a.seed1;
  if (a.m_off) {
    r = find_item(a.m_off, a.count, t);
    m0 = a.m_off[r];
    M = a.m_off[r + 1] - m0;
    ls = a.len_scale[r];
    seed = a.seeds[r];
  }
  const int64_t m = t - m0;
  uint64_t w[8];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    uint64_t c[4] = {static_cast<uint64_t>(2 * m + b + 1), 0, 0, 0};
    philox4x64_10(c, seed, a.stream_id);
#pragma unroll
    for (int i = 0; i < 4; ++i) w[4 * b + i] = c[i];
  }
  double n[8];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double u0 = static_cast<double>(w[2 * i] >> 11) * 0x1.0p-53;
    const double u1 = static_cast<double>(w[2 * i + 1] >> 11) * 0x1.0p-53;
    const double rho = sqrt(-2.0 * log(1.0 - u0));
    double s, cs;
    sincospi(2.0 * u1, &s, &cs);
    n[2 * i] = rho * cs;
    n[2 * i + 1] = rho * s;
  }
  const double scale = a.model == 1 ? 1.0 / (ls * fmax(fabs(n[6]), 0x1.0p-1022)) : 1.0 / ls;
  double* cov = a.cov + a.dim * m0;
  for (int d = 0; d < a.dim; ++d) cov[d * M + m] = n[d] * scale;
  a.z1[m0 + m] = n[4];
  a.z2[m0 + m] = n[5];
}

}  // namespace

hipError_t launch_sample_modes(const SampleArgs& a, hipStream_t stream) {
  if (a.m_total <= 0) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((a.m_total + 255) / 256);
  k_sample_modes<<<blocks, 256, 0, stream>>>(a);
  return hipGetLastError();
}

}  // namespace gsc
