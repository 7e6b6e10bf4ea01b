// HIP kernels for gfx950: Wilson hop stencil (4D and 5D / multi-RHS), fifth-dimension
// operators, BLAS1, deterministic reductions and the device-side CG scalar updates.
//
// Semantics (paths relative to the Grid source tree):
//   hop: out(x) = sum over 8 legs of recon(U * proj(in(x +- mu))), links pre-scaled by -1/2;
//        DaggerNo: leg at x-mu with (1+gamma_mu) and U_mu(x-mu)^dag, leg at x+mu with (1-gamma_mu)
//        and U_mu(x); DaggerYes swaps the projectors.
//        Grid/qcd/action/fermion/implementation/WilsonKernelsImplementation.h:112-163,
//        Grid/qcd/spin/TwoSpinor.h:67-323, Grid/qcd/action/fermion/WilsonImpl.h:83-101.
//   5D:  links shared across s (WilsonFermion5DImplementation.h:414-452).
//   fifth-dimension operators (M5D family, Mooee, MooeeInv) are real Ls x Ls matrices per
//   chirality (CayleyFermion5Dcache.h:41-230); the kernel applies them from LDS.
//   reductions: per-site products in the field precision, sums in double
//   (Grid/lattice/Lattice_reduction.h:255-287), in a fixed order.
#include <climits>

#include "gx_internal.h"

namespace gx {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T r, i;
};

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return {a.r + b.r, a.i + b.i}; }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return {a.r - b.r, a.i - b.i}; }
// a + i b, a - i b
template <typename T>
__device__ __forceinline__ cx<T> caddi(cx<T> a, cx<T> b) { return {a.r - b.i, a.i + b.r}; }
template <typename T>
__device__ __forceinline__ cx<T> csubi(cx<T> a, cx<T> b) { return {a.r + b.i, a.i - b.r}; }

struct KGeom {
  int L[4];
  int Lxh;
  int64_t H4;
  int64_t stride;          // complex per component per checkerboard = H4*Ls
  int64_t sY, sZ, sT;      // half-site strides of y, z, t
};

static KGeom kgeom(const Geom &g) {
  KGeom k;
  for (int d = 0; d < 4; d++) k.L[d] = g.L[d];
  k.Lxh = g.L[0] / 2;
  k.H4 = g.H4;
  k.stride = g.H4 * g.Ls;
  k.sY = k.Lxh;
  k.sZ = k.sY * g.L[1];
  k.sT = k.sZ * g.L[2];
  return k;
}

__device__ __forceinline__ bool gated(const CGState *gate) {
  return gate && (gate->k > gate->k_conv || gate->k > gate->kmax);
}

// spin projection (1 + gamma_mu) (PLUS) or (1 - gamma_mu) on a 12-component spinor f[spin*3+c],
// giving the 2-spinor h[spin*3+c], spin in {0,1}.  TwoSpinor.h:75-133
template <typename T, int MU, bool PLUS>
__device__ __forceinline__ void project(const cx<T> *f, cx<T> *h) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<T> f0 = f[c], f1 = f[3 + c], f2 = f[6 + c], f3 = f[9 + c];
    if (MU == 0) {
      if (PLUS) { h[c] = caddi(f0, f3); h[3 + c] = caddi(f1, f2); }
      else      { h[c] = csubi(f0, f3); h[3 + c] = csubi(f1, f2); }
    } else if (MU == 1) {
      if (PLUS) { h[c] = csub(f0, f3); h[3 + c] = cadd(f1, f2); }
      else      { h[c] = cadd(f0, f3); h[3 + c] = csub(f1, f2); }
    } else if (MU == 2) {
      if (PLUS) { h[c] = caddi(f0, f2); h[3 + c] = csubi(f1, f3); }
      else      { h[c] = csubi(f0, f2); h[3 + c] = caddi(f1, f3); }
    } else {
      if (PLUS) { h[c] = cadd(f0, f2); h[3 + c] = cadd(f1, f3); }
      else      { h[c] = csub(f0, f2); h[3 + c] = csub(f1, f3); }
    }
  }
}

// accumulate the reconstruction of a projected 2-spinor, TwoSpinor.h:193-323
template <typename T, int MU, bool PLUS>
__device__ __forceinline__ void accum_recon(cx<T> *a, const cx<T> *g) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<T> h0 = g[c], h1 = g[3 + c];
    a[c] = cadd(a[c], h0);
    a[3 + c] = cadd(a[3 + c], h1);
    if (MU == 0) {
      if (PLUS) { a[6 + c] = csubi(a[6 + c], h1); a[9 + c] = csubi(a[9 + c], h0); }
      else      { a[6 + c] = caddi(a[6 + c], h1); a[9 + c] = caddi(a[9 + c], h0); }
    } else if (MU == 1) {
      if (PLUS) { a[6 + c] = cadd(a[6 + c], h1); a[9 + c] = csub(a[9 + c], h0); }
      else      { a[6 + c] = csub(a[6 + c], h1); a[9 + c] = cadd(a[9 + c], h0); }
    } else if (MU == 2) {
      if (PLUS) { a[6 + c] = csubi(a[6 + c], h0); a[9 + c] = caddi(a[9 + c], h1); }
      else      { a[6 + c] = caddi(a[6 + c], h0); a[9 + c] = csubi(a[9 + c], h1); }
    } else {
      if (PLUS) { a[6 + c] = cadd(a[6 + c], h0); a[9 + c] = cadd(a[9 + c], h1); }
      else      { a[6 + c] = csub(a[6 + c], h0); a[9 + c] = csub(a[9 + c], h1); }
    }
  }
}

// g = U h for both spin rows (SU(3) x colour vector, twice)
template <typename T>
__device__ __forceinline__ void su3_mul(const cx<T> *u, const cx<T> *h, cx<T> *g) {
#pragma unroll
  for (int sp = 0; sp < 2; sp++)
#pragma unroll
    for (int a = 0; a < 3; a++) {
      T re = 0, im = 0;
#pragma unroll
      for (int b = 0; b < 3; b++) {
        const cx<T> uu = u[a * 3 + b], hh = h[sp * 3 + b];
        re = fma(uu.r, hh.r, re);
        re = fma(-uu.i, hh.i, re);
        im = fma(uu.r, hh.i, im);
        im = fma(uu.i, hh.r, im);
      }
      g[sp * 3 + a] = {re, im};
    }
}

template <typename T, int LS, int P, bool PLUS>
__device__ __forceinline__ void hop_leg(cx<T> *acc, const cx<T> *__restrict__ in, const cx<T> *__restrict__ U,
                                        int64_t nb, int s, int64_t h, const KGeom &g) {
  constexpr int MU = P & 3;
  cx<T> f[12];
  const int64_t idx = nb * LS + s;
#pragma unroll
  for (int k = 0; k < 12; k++) f[k] = in[k * g.stride + idx];
  cx<T> hs[6];
  project<T, MU, PLUS>(f, hs);
  cx<T> u[9];
#pragma unroll
  for (int k = 0; k < 9; k++) u[k] = U[(P * 9 + k) * g.H4 + h];
  cx<T> gs[6];
  su3_mul<T>(u, hs, gs);
  accum_recon<T, MU, PLUS>(acc, gs);
}

// One lane per (4D half-site h, s); s fastest, 64/LS sites per wave; lanes of one site share
// the link loads (same address within the wave).
template <typename T, int LS, bool DAG>
__global__ void __launch_bounds__(256) k_dhop(const cx<T> *__restrict__ in, cx<T> *__restrict__ out,
                                              const cx<T> *__restrict__ U, KGeom g, int cb_out,
                                              const CGState *__restrict__ gate) {
  if (gated(gate)) return;
  constexpr int SPW = 64 / LS;
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int ls = lane / LS;
  const int s = lane - ls * LS;
  if (ls >= SPW) return;
  const int64_t h = wave * SPW + ls;
  if (h >= g.H4) return;
  int64_t r = h;
  const int xh = (int)(r % g.Lxh);
  r /= g.Lxh;
  const int y = (int)(r % g.L[1]);
  r /= g.L[1];
  const int z = (int)(r % g.L[2]);
  const int t = (int)(r / g.L[2]);
  const int x = 2 * xh + ((cb_out + y + z + t) & 1);
  const int64_t base = h - xh;
  const int xp = (x + 1 == g.L[0]) ? 0 : x + 1;
  const int xm = (x == 0) ? g.L[0] - 1 : x - 1;
  int64_t nb[8];
  nb[0] = base + (xp >> 1);
  nb[4] = base + (xm >> 1);
  nb[1] = (y + 1 == g.L[1]) ? h - (int64_t)(g.L[1] - 1) * g.sY : h + g.sY;
  nb[5] = (y == 0) ? h + (int64_t)(g.L[1] - 1) * g.sY : h - g.sY;
  nb[2] = (z + 1 == g.L[2]) ? h - (int64_t)(g.L[2] - 1) * g.sZ : h + g.sZ;
  nb[6] = (z == 0) ? h + (int64_t)(g.L[2] - 1) * g.sZ : h - g.sZ;
  nb[3] = (t + 1 == g.L[3]) ? h - (int64_t)(g.L[3] - 1) * g.sT : h + g.sT;
  nb[7] = (t == 0) ? h + (int64_t)(g.L[3] - 1) * g.sT : h - g.sT;

  cx<T> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0, 0};
  if (!DAG) {
    hop_leg<T, LS, 4, true>(acc, in, U, nb[4], s, h, g);
    hop_leg<T, LS, 5, true>(acc, in, U, nb[5], s, h, g);
    hop_leg<T, LS, 6, true>(acc, in, U, nb[6], s, h, g);
    hop_leg<T, LS, 7, true>(acc, in, U, nb[7], s, h, g);
    hop_leg<T, LS, 0, false>(acc, in, U, nb[0], s, h, g);
    hop_leg<T, LS, 1, false>(acc, in, U, nb[1], s, h, g);
    hop_leg<T, LS, 2, false>(acc, in, U, nb[2], s, h, g);
    hop_leg<T, LS, 3, false>(acc, in, U, nb[3], s, h, g);
  } else {
    hop_leg<T, LS, 0, true>(acc, in, U, nb[0], s, h, g);
    hop_leg<T, LS, 1, true>(acc, in, U, nb[1], s, h, g);
    hop_leg<T, LS, 2, true>(acc, in, U, nb[2], s, h, g);
    hop_leg<T, LS, 3, true>(acc, in, U, nb[3], s, h, g);
    hop_leg<T, LS, 4, false>(acc, in, U, nb[4], s, h, g);
    hop_leg<T, LS, 5, false>(acc, in, U, nb[5], s, h, g);
    hop_leg<T, LS, 6, false>(acc, in, U, nb[6], s, h, g);
    hop_leg<T, LS, 7, false>(acc, in, U, nb[7], s, h, g);
  }
  const int64_t o = h * LS + s;
#pragma unroll
  for (int k = 0; k < 12; k++) out[k * g.stride + o] = acc[k];
}

#define GX_LS_CASES(M) \
  M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12) M(13) M(14) M(15) M(16)

template <typename T>
int launch_dhop(const Geom &g, const void *in, void *out, const void *U, int cb_out, int dag, const CGState *gate,
                hipStream_t st) {
  const KGeom k = kgeom(g);
  const int spw = 64 / g.Ls;
  const int64_t waves = (g.H4 + spw - 1) / spw;
  const int64_t blocks = (waves * 64 + 255) / 256;
  const cx<T> *pin = (const cx<T> *)in;
  cx<T> *pout = (cx<T> *)out;
  const cx<T> *pU = (const cx<T> *)U;
#define GX_DHOP_CASE(LS)                                                                          \
  case LS:                                                                                        \
    if (dag)                                                                                      \
      hipLaunchKernelGGL((k_dhop<T, LS, true>), dim3(blocks), dim3(256), 0, st, pin, pout, pU, k, cb_out, gate); \
    else                                                                                          \
      hipLaunchKernelGGL((k_dhop<T, LS, false>), dim3(blocks), dim3(256), 0, st, pin, pout, pU, k, cb_out, gate); \
    break;
  switch (g.Ls) {
    GX_LS_CASES(GX_DHOP_CASE)
    default:
      set_error("dhop: Ls=%d unsupported (1..%d)", g.Ls, kMaxLs);
      return GX_EINVAL;
  }
#undef GX_DHOP_CASE
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// out_s = sum_t A_chi[s][t] in_t  (+ phi_s).  mats: [2][LS][LS] in T, chirality 0 = spins 0,1.
template <typename T, int LS>
__global__ void __launch_bounds__(256) k_smat(const cx<T> *__restrict__ in, const cx<T> *__restrict__ phi,
                                              cx<T> *__restrict__ out, const T *__restrict__ mats, int64_t H4,
                                              int64_t stride, const CGState *__restrict__ gate) {
  if (gated(gate)) return;
  constexpr int SPW = 64 / LS;
  __shared__ cx<T> lds[4][12][64];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int ls = lane / LS;
  const int s = lane - ls * LS;
  const int64_t h = wave * SPW + ls;
  const bool valid = (ls < SPW) && (h < H4);
  const int64_t o = h * LS + s;
#pragma unroll
  for (int k = 0; k < 12; k++) lds[w][k][lane] = valid ? in[k * stride + o] : cx<T>{0, 0};
  __syncthreads();
  if (!valid) return;
  T ap[LS], am[LS];
#pragma unroll
  for (int t = 0; t < LS; t++) {
    ap[t] = mats[s * LS + t];
    am[t] = mats[LS * LS + s * LS + t];
  }
#pragma unroll
  for (int k = 0; k < 12; k++) {
    T re = 0, im = 0;
#pragma unroll
    for (int t = 0; t < LS; t++) {
      const cx<T> v = lds[w][k][ls * LS + t];
      const T a = (k < 6) ? ap[t] : am[t];
      re = fma(a, v.r, re);
      im = fma(a, v.i, im);
    }
    if (phi) {
      const cx<T> p = phi[k * stride + o];
      re += p.r;
      im += p.i;
    }
    out[k * stride + o] = {re, im};
  }
}

template <typename T>
int launch_smat(const Geom &g, const void *in, const void *phi, void *out, const void *mats, const CGState *gate,
                hipStream_t st) {
  const int spw = 64 / g.Ls;
  const int64_t waves = (g.H4 + spw - 1) / spw;
  const int64_t blocks = (waves * 64 + 255) / 256;
  const int64_t stride = g.H4 * g.Ls;
  const cx<T> *pin = (const cx<T> *)in;
  const cx<T> *pphi = (const cx<T> *)phi;
  cx<T> *pout = (cx<T> *)out;
  const T *pm = (const T *)mats;
#define GX_SMAT_CASE(LS)                                                                                   \
  case LS:                                                                                                 \
    hipLaunchKernelGGL((k_smat<T, LS>), dim3(blocks), dim3(256), 0, st, pin, pphi, pout, pm, g.H4, stride, gate); \
    break;
  switch (g.Ls) {
    GX_LS_CASES(GX_SMAT_CASE)
    default:
      set_error("smat: Ls=%d unsupported", g.Ls);
      return GX_EINVAL;
  }
#undef GX_SMAT_CASE
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- BLAS1 ----------------
template <typename T>
__global__ void __launch_bounds__(256) k_axpby(int64_t n, T a, const cx<T> *__restrict__ x, T b,
                                               const cx<T> *__restrict__ y, cx<T> *__restrict__ z,
                                               const CGState *__restrict__ gate) {
  if (gated(gate)) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cx<T> xv = x[i];
    if (y) {
      const cx<T> yv = y[i];
      z[i] = {fma(a, xv.r, b * yv.r), fma(a, xv.i, b * yv.i)};
    } else {
      z[i] = {a * xv.r, a * xv.i};
    }
  }
}

static int64_t stream_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

template <typename T>
int launch_axpby(int64_t n, double a, const void *x, double b, const void *y, void *z, const CGState *gate,
                 hipStream_t st) {
  hipLaunchKernelGGL((k_axpby<T>), dim3(stream_blocks(n)), dim3(256), 0, st, n, (T)a, (const cx<T> *)x, (T)b,
                     (const cx<T> *)y, (cx<T> *)z, gate);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

template <typename Tin, typename Tout>
__global__ void __launch_bounds__(256) k_convert(int64_t n, const Tin *__restrict__ in, Tout *__restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = (Tout)in[i];
}

template <typename Tin, typename Tout>
int launch_convert(int64_t n, const void *in, void *out, hipStream_t st) {
  hipLaunchKernelGGL((k_convert<Tin, Tout>), dim3(stream_blocks(n)), dim3(256), 0, st, n, (const Tin *)in,
                     (Tout *)out);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- deterministic reductions ----------------
// Block tree (wave shuffles, then LDS across the 4 waves) followed by a last-block-done
// ticket: the last block sums the per-block partials in block order.  Grid size depends only
// on the problem size, so the result is bitwise reproducible.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double (*sh)[4]) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int j = 0; j < NV; j++) v[j] = wave_sum(v[j]);
  if (lane == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) sh[j][w] = v[j];
  __syncthreads();
#pragma unroll
  for (int j = 0; j < NV; j++) v[j] = (sh[j][0] + sh[j][1]) + (sh[j][2] + sh[j][3]);
  __syncthreads();
}

// Returns true in every thread of the last block to finish; partial[blockIdx*NV + j] = v[j].
template <int NV>
__device__ __forceinline__ bool publish_partial(const double (&v)[NV], double *partial, unsigned int *ticket) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int j = 0; j < NV; j++) partial[blockIdx.x * NV + j] = v[j];
    __threadfence();
    const unsigned int t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  return last;
}

template <int NV>
__device__ __forceinline__ void final_sum(double (&v)[NV], const double *partial, unsigned int *ticket,
                                          double (*sh)[4]) {
  __threadfence();
#pragma unroll
  for (int j = 0; j < NV; j++) v[j] = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x)
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] += ((volatile const double *)partial)[b * NV + j];
  block_sum<NV>(v, sh);
  if (threadIdx.x == 0) *ticket = 0;
}

// site j of a (possibly full) field: half = j / stride, i = j % stride;
// component k at half*12*stride + k*stride + i.
template <typename T, int MODE>
__global__ void __launch_bounds__(256) k_reduce(int64_t nsites, int64_t stride, const cx<T> *__restrict__ x,
                                                const cx<T> *__restrict__ y, cx<T> *__restrict__ z, T a,
                                                double *partial, unsigned int *ticket, double *result) {
  __shared__ double sh[2][4];
  double v[2] = {0, 0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride;
    const int64_t base = hf * 12 * stride + (j - hf * stride);
    T sr = 0, si = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const cx<T> xv = x[base + k * stride];
      if (MODE == 0) {
        const cx<T> yv = y[base + k * stride];
        sr = fma(xv.r, yv.r, sr);
        sr = fma(xv.i, yv.i, sr);
        si = fma(xv.r, yv.i, si);
        si = fma(-xv.i, yv.r, si);
      } else if (MODE == 1) {
        sr = fma(xv.r, xv.r, sr);
        sr = fma(xv.i, xv.i, sr);
      } else {
        const cx<T> yv = y[base + k * stride];
        const cx<T> zv = {fma(a, xv.r, yv.r), fma(a, xv.i, yv.i)};
        z[base + k * stride] = zv;
        sr = fma(zv.r, zv.r, sr);
        sr = fma(zv.i, zv.i, sr);
      }
    }
    v[0] += (double)sr;
    v[1] += (double)si;
  }
  block_sum<2>(v, sh);
  if (publish_partial<2>(v, partial, ticket)) {
    final_sum<2>(v, partial, ticket, sh);
    if (threadIdx.x == 0) {
      result[0] = v[0];
      if (MODE == 0) result[1] = v[1];
    }
  }
}

static int red_blocks(int64_t nsites) {
  int64_t b = (nsites + 255) / 256;
  if (b > kRedBlocks) b = kRedBlocks;
  return (int)(b < 1 ? 1 : b);
}

template <typename T>
int launch_reduce(Ctx *ctx, int mode, int64_t nsites, int64_t stride, const void *x, const void *y, void *z, double a,
                  int slot) {
  const int nb = red_blocks(nsites);
  const cx<T> *px = (const cx<T> *)x;
  const cx<T> *py = (const cx<T> *)y;
  cx<T> *pz = (cx<T> *)z;
  double *res = ctx->red_result + slot;
  if (mode == 0)
    hipLaunchKernelGGL((k_reduce<T, 0>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial, ctx->red_ticket, res);
  else if (mode == 1)
    hipLaunchKernelGGL((k_reduce<T, 1>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial, ctx->red_ticket, res);
  else
    hipLaunchKernelGGL((k_reduce<T, 2>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial, ctx->red_ticket, res);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- CG iteration on device scalars (ConjugateGradient.h:151-198) ----------------
// d = Re<p|mmp>; a = c / d
template <typename T>
__global__ void __launch_bounds__(256) k_cg_dot(int64_t nsites, int64_t stride, const cx<T> *__restrict__ p,
                                                const cx<T> *__restrict__ mmp, CGState *st, double *partial,
                                                unsigned int *ticket) {
  if (gated(st)) return;
  __shared__ double sh[1][4];
  double v[1] = {0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride;
    const int64_t base = hf * 12 * stride + (j - hf * stride);
    T sr = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const cx<T> xv = p[base + k * stride], yv = mmp[base + k * stride];
      sr = fma(xv.r, yv.r, sr);
      sr = fma(xv.i, yv.i, sr);
    }
    v[0] += (double)sr;
  }
  block_sum<1>(v, sh);
  if (publish_partial<1>(v, partial, ticket)) {
    final_sum<1>(v, partial, ticket, sh);
    if (threadIdx.x == 0) {
      st->d = v[0];
      st->a = st->c / st->d;
    }
  }
}

// r = r - a mmp; cp = |r|^2; b = cp / c; convergence test cp <= rsq
template <typename T>
__global__ void __launch_bounds__(256) k_cg_axpy_norm(int64_t nsites, int64_t stride, cx<T> *__restrict__ r,
                                                      const cx<T> *__restrict__ mmp, CGState *st, double *partial,
                                                      unsigned int *ticket) {
  if (gated(st)) return;
  __shared__ double sh[1][4];
  const T a = (T)(-st->a);
  double v[1] = {0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride;
    const int64_t base = hf * 12 * stride + (j - hf * stride);
    T sr = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const cx<T> xv = mmp[base + k * stride], yv = r[base + k * stride];
      const cx<T> zv = {fma(a, xv.r, yv.r), fma(a, xv.i, yv.i)};
      r[base + k * stride] = zv;
      sr = fma(zv.r, zv.r, sr);
      sr = fma(zv.i, zv.i, sr);
    }
    v[0] += (double)sr;
  }
  block_sum<1>(v, sh);
  if (publish_partial<1>(v, partial, ticket)) {
    final_sum<1>(v, partial, ticket, sh);
    if (threadIdx.x == 0) {
      const double cp = v[0];
      st->b = cp / st->c;
      st->cp = cp;
      st->c = cp;
      if (cp <= st->rsq && st->k_conv == INT_MAX) st->k_conv = st->k;
    }
  }
}

// psi = a p + psi; p = b p + r; the last block advances the iteration counter
template <typename T>
__global__ void __launch_bounds__(256) k_cg_update(int64_t n, cx<T> *__restrict__ psi, cx<T> *__restrict__ p,
                                                   const cx<T> *__restrict__ r, CGState *st, unsigned int *ticket) {
  if (gated(st)) return;
  const T a = (T)st->a, b = (T)st->b;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cx<T> pv = p[i], xv = psi[i], rv = r[i];
    psi[i] = {fma(a, pv.r, xv.r), fma(a, pv.i, xv.i)};
    p[i] = {fma(b, pv.r, rv.r), fma(b, pv.i, rv.i)};
  }
  __shared__ bool last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned int t = atomicAdd(ticket, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (last && threadIdx.x == 0) {
    *ticket = 0;
    st->k = st->k + 1;
  }
}

template <typename T>
int launch_cg_dot(Ctx *ctx, int64_t nsites, int64_t stride, const void *p, const void *mmp, CGState *state) {
  hipLaunchKernelGGL((k_cg_dot<T>), dim3(red_blocks(nsites)), dim3(256), 0, ctx->stream, nsites, stride,
                     (const cx<T> *)p, (const cx<T> *)mmp, state, ctx->red_partial, ctx->red_ticket);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}
template <typename T>
int launch_cg_axpy_norm(Ctx *ctx, int64_t nsites, int64_t stride, void *r, const void *mmp, CGState *state) {
  hipLaunchKernelGGL((k_cg_axpy_norm<T>), dim3(red_blocks(nsites)), dim3(256), 0, ctx->stream, nsites, stride,
                     (cx<T> *)r, (const cx<T> *)mmp, state, ctx->red_partial, ctx->red_ticket);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}
template <typename T>
int launch_cg_update(Ctx *ctx, int64_t n, void *psi, void *p, const void *r, CGState *state) {
  int64_t nb = stream_blocks(n);
  if (nb > 4096) nb = 4096;
  hipLaunchKernelGGL((k_cg_update<T>), dim3(nb), dim3(256), 0, ctx->stream, n, (cx<T> *)psi, (cx<T> *)p,
                     (const cx<T> *)r, state, ctx->red_ticket + 1);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

#define GX_INST(T)                                                                                                 \
  template int launch_dhop<T>(const Geom &, const void *, void *, const void *, int, int, const CGState *,        \
                              hipStream_t);                                                                        \
  template int launch_smat<T>(const Geom &, const void *, const void *, void *, const void *, const CGState *,     \
                              hipStream_t);                                                                        \
  template int launch_axpby<T>(int64_t, double, const void *, double, const void *, void *, const CGState *,       \
                               hipStream_t);                                                                       \
  template int launch_reduce<T>(Ctx *, int, int64_t, int64_t, const void *, const void *, void *, double, int);   \
  template int launch_cg_dot<T>(Ctx *, int64_t, int64_t, const void *, const void *, CGState *);                   \
  template int launch_cg_axpy_norm<T>(Ctx *, int64_t, int64_t, void *, const void *, CGState *);                   \
  template int launch_cg_update<T>(Ctx *, int64_t, void *, void *, const void *, CGState *);
GX_INST(double)
GX_INST(float)
template int launch_convert<double, float>(int64_t, const void *, void *, hipStream_t);
template int launch_convert<float, double>(int64_t, const void *, void *, hipStream_t);

}  // namespace gx
