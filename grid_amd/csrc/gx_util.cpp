// Synthetic inputs: random SU(3) gauge fields U = exp(i eps H) and Gaussian / point sources.
// Gaussian numbers come from SplitMix64 counters keyed by (seed, global index) and the
// Box-Muller transform, so every layout and device sees the same field for the same seed.
// Grid's own recipe (GaugeGroup.h:331-403, gaussian -> Ta -> taExp) is replaced by the
// exp(i eps H) prescription of the benchmark configurations.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/gx.h"

namespace {

typedef std::complex<double> cd;

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

// uniform in (0, 1)
inline double u01(uint64_t key, uint64_t ctr) {
  const uint64_t r = splitmix64(key ^ splitmix64(ctr));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

inline cd gauss(uint64_t key, uint64_t ctr) {
  const double a = u01(key, 2 * ctr), b = u01(key, 2 * ctr + 1);
  const double r = std::sqrt(-2.0 * std::log(a));
  const double t = 2.0 * M_PI * b;
  return cd(r * std::cos(t), r * std::sin(t));
}

struct M3 {
  cd m[3][3];
};

M3 mul(const M3 &a, const M3 &b) {
  M3 r;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      cd s = 0;
      for (int k = 0; k < 3; k++) s += a.m[i][k] * b.m[k][j];
      r.m[i][j] = s;
    }
  return r;
}

// exp(A) by scaling and squaring of a degree-16 Taylor series (Horner form).
M3 expm(const M3 &A) {
  double nrm = 0;
  for (int i = 0; i < 3; i++) {
    double s = 0;
    for (int j = 0; j < 3; j++) s += std::abs(A.m[i][j]);
    nrm = std::max(nrm, s);
  }
  int sq = 0;
  while (nrm > 0.25) {
    nrm *= 0.5;
    sq++;
  }
  M3 B = A;
  const double f = std::ldexp(1.0, -sq);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) B.m[i][j] *= f;
  M3 R;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R.m[i][j] = (i == j) ? 1.0 : 0.0;
  for (int n = 16; n >= 1; n--) {
    M3 T = mul(B, R);
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) R.m[i][j] = ((i == j) ? 1.0 : 0.0) + T.m[i][j] / (double)n;
  }
  for (int k = 0; k < sq; k++) R = mul(R, R);
  return R;
}

// Project onto SU(3): Gram-Schmidt on rows 0 and 1, row 2 = conj(row0 x row1).
void project_su3(M3 &U) {
  cd *r0 = U.m[0], *r1 = U.m[1], *r2 = U.m[2];
  double n0 = std::sqrt(std::norm(r0[0]) + std::norm(r0[1]) + std::norm(r0[2]));
  for (int i = 0; i < 3; i++) r0[i] /= n0;
  cd d = std::conj(r0[0]) * r1[0] + std::conj(r0[1]) * r1[1] + std::conj(r0[2]) * r1[2];
  for (int i = 0; i < 3; i++) r1[i] -= d * r0[i];
  double n1 = std::sqrt(std::norm(r1[0]) + std::norm(r1[1]) + std::norm(r1[2]));
  for (int i = 0; i < 3; i++) r1[i] /= n1;
  r2[0] = std::conj(r0[1] * r1[2] - r0[2] * r1[1]);
  r2[1] = std::conj(r0[2] * r1[0] - r0[0] * r1[2]);
  r2[2] = std::conj(r0[0] * r1[1] - r0[1] * r1[0]);
}

template <typename F>
void par(int64_t n, F f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 32) nt = 32;
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=]() { f(b, e); });
  }
  for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

int gx_random_gauge(const int dims[4], uint64_t seed, double eps, double *U) {
  if (!dims || !U) return GX_EINVAL;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const uint64_t key = splitmix64(seed * 0x2545F4914F6CDD1DULL + 0x6A09E667F3BCC909ULL);
  par(V * 4, [&](int64_t b, int64_t e) {
    for (int64_t l = b; l < e; l++) {
      M3 G, H, A;
      for (int j = 0; j < 9; j++) G.m[j / 3][j % 3] = gauss(key, (uint64_t)l * 9 + j);
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H.m[i][j] = 0.5 * (G.m[i][j] + std::conj(G.m[j][i]));
      const double tr = (H.m[0][0].real() + H.m[1][1].real() + H.m[2][2].real()) / 3.0;
      for (int i = 0; i < 3; i++) H.m[i][i] -= tr;
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) A.m[i][j] = cd(0, eps) * H.m[i][j];
      M3 R = expm(A);
      project_su3(R);
      double *o = U + l * 18;
      for (int j = 0; j < 9; j++) {
        o[2 * j] = R.m[j / 3][j % 3].real();
        o[2 * j + 1] = R.m[j / 3][j % 3].imag();
      }
    }
  });
  return GX_OK;
}

int gx_gaussian_fermion(const int dims[4], int Ls, uint64_t seed, double *psi) {
  if (!dims || !psi || Ls < 1) return GX_EINVAL;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const uint64_t key = splitmix64(seed * 0x9E6C63D0676A9A99ULL + 0xBB67AE8584CAA73BULL);
  par(V * Ls, [&](int64_t b, int64_t e) {
    for (int64_t i = b; i < e; i++)
      for (int k = 0; k < 12; k++) {
        const cd g = gauss(key, (uint64_t)i * 12 + k);
        psi[(i * 12 + k) * 2] = g.real();
        psi[(i * 12 + k) * 2 + 1] = g.imag();
      }
  });
  return GX_OK;
}

int gx_point_source(const int dims[4], int Ls, const int x[4], int spin, int colour, double *psi) {
  if (!dims || !x || !psi || Ls < 1 || spin < 0 || spin > 3 || colour < 0 || colour > 2) return GX_EINVAL;
  for (int d = 0; d < 4; d++)
    if (x[d] < 0 || x[d] >= dims[d]) return GX_EINVAL;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  std::memset(psi, 0, sizeof(double) * 2 * 12 * V * Ls);
  const int64_t site = x[0] + (int64_t)dims[0] * (x[1] + (int64_t)dims[1] * (x[2] + (int64_t)dims[2] * x[3]));
  for (int s = 0; s < Ls; s++) psi[((site * Ls + s) * 12 + spin * 3 + colour) * 2] = 1.0;
  return GX_OK;
}

}  // extern "C"
