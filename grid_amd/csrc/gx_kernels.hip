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
#include <cfloat>
#include <climits>
#include <cstring>
#include <type_traits>

#include "gx_internal.h"

// The file is compiled as three translation units (GX_PART 0, 1, 2 instantiate the FP64, FP32 and
// FP16 storage types; gx_kernels_f32.hip and gx_kernels_f16.hip include it) so the storage types
// build in parallel.  Non-template kernels are static: every unit carries its own copy.
#ifndef GX_PART
#define GX_PART 0
#endif

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

// Storage of field entries: double, float, or FP16 (half_t).  FP16 keeps a complex number as two
// binary16 values and every (h, s) site of a half block carries a float scale in column 12 (its
// largest |component|), so stored values lie in [-1, 1]; its arithmetic type is float.  FP16 links
// are stored unscaled (SU(3) entries are bounded by 1).  VW complex numbers fill 16 bytes: a site's
// components are stored in 16-byte groups of VW consecutive components, so a lane moves a site with
// 12 / VW loads or stores of 16 bytes whatever the precision.
template <typename S>
struct Sto {
  typedef S C;
  static constexpr bool H = false;
  static constexpr int NCOL = 12;
  static constexpr int VW = 16 / (int)sizeof(cx<S>);
};
template <>
struct Sto<half_t> {
  typedef float C;
  static constexpr bool H = true;
  static constexpr int NCOL = 13;
  static constexpr int VW = 4;
};

// Position of component k < 12 of entry e in a half field of storage S: blocks of 64 consecutive
// entries; within a block, group g = k / VW of all 64 entries, entry by entry, each entry's group
// the VW components VW g .. VW g + VW - 1:  (e / 64) NCOL 64 + (k / VW) 64 VW + (e % 64) VW + k % VW.
// A wave of 64 consecutive entries moves one contiguous 1 KB run per group.
template <typename S, typename I>
__device__ __forceinline__ I bidx(I e, int k) {
  constexpr int VW = Sto<S>::VW;
  return (e >> 6) * (I)(Sto<S>::NCOL * kBlk) + (I)((k / VW) * kBlk * VW) + (e & (I)(kBlk - 1)) * (I)VW + (I)(k % VW);
}
// FP16 site scale of entry e: a float in column 12 of its block
template <typename S, typename I>
__device__ __forceinline__ I sidx(I e) {
  return (e >> 6) * (I)(Sto<S>::NCOL * kBlk) + (I)(12 * kBlk) + (e & (I)(kBlk - 1));
}
// entry of stored position i of an FP64 or FP32 half field (the inverse of bidx over the block)
template <typename S>
__device__ __forceinline__ int64_t entry_of(int64_t i) {
  constexpr int VW = Sto<S>::VW;
  const int64_t blk = i / (12 * kBlk), o = i - blk * 12 * kBlk;
  return blk * kBlk + (o % (kBlk * VW)) / VW;
}

template <typename S>
__device__ __forceinline__ cx<typename Sto<S>::C> ldc(const cx<S> *__restrict__ p, int64_t i) {
  const cx<S> v = p[i];
  return {(typename Sto<S>::C)v.r, (typename Sto<S>::C)v.i};
}

// scale of entry i of the half block at p (FP16 storage; 1 otherwise)
template <typename S, typename I>
__device__ __forceinline__ typename Sto<S>::C site_scale(const cx<S> *__restrict__ p, I i) {
  if constexpr (Sto<S>::H) return *reinterpret_cast<const float *>(p + sidx<S, I>(i));
  else return (typename Sto<S>::C)1;
}

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// The VW components of group g of the entry whose group 0 is at q, as values of the arithmetic
// type (FP16 unscaled).
template <typename S>
__device__ __forceinline__ void load_group(cx<typename Sto<S>::C> *f, const cx<S> *__restrict__ q, int g) {
  constexpr int VW = Sto<S>::VW;
  const cx<S> *__restrict__ a = q + g * kBlk * VW;
  if constexpr (VW == 1) {
    const f64x2 v = *reinterpret_cast<const f64x2 *>(a);
    f[0] = {v.x, v.y};
  } else if constexpr (!Sto<S>::H) {
    const f32x4 v = *reinterpret_cast<const f32x4 *>(a);
    f[0] = {v.x, v.y};
    f[1] = {v.z, v.w};
  } else {
    const f16x8 v = *reinterpret_cast<const f16x8 *>(a);
#pragma unroll
    for (int c = 0; c < 4; c++) f[c] = {(float)v[2 * c], (float)v[2 * c + 1]};
  }
}

// a * b rounded on its own: never contracted into an FMA with the consumer of the product
__device__ __forceinline__ float mul_rounded(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}

// The 12 components of entry i (FP16: multiplied by the site's scale), 12 / VW loads of 16 bytes.
template <typename S, typename I>
__device__ __forceinline__ void load_site(cx<typename Sto<S>::C> *f, const cx<S> *__restrict__ p, I i) {
  constexpr int VW = Sto<S>::VW;
  const cx<S> *__restrict__ q = p + bidx<S, I>(i, 0);
#pragma unroll
  for (int g = 0; g < 12 / VW; g++) load_group<S>(f + g * VW, q, g);
  if constexpr (Sto<S>::H) {
    // rounded products (no contraction into the consumer's FMAs): a site's values are the same
    // whether a kernel uses them at once or stages them in LDS first
    const float sc = *reinterpret_cast<const float *>(p + sidx<S, I>(i));
#pragma unroll
    for (int k = 0; k < 12; k++) f[k] = {mul_rounded(f[k].r, sc), mul_rounded(f[k].i, sc)};
  }
}

// The 12 components of entry i as stored, and the factor they stand for: FP16 storage returns
// the site's scale unapplied (the hop folds it into its reconstruction, 24 multiplies fewer per
// leg), other storages 1.
template <typename S, typename I>
__device__ __forceinline__ typename Sto<S>::C load_site_raw(cx<typename Sto<S>::C> *f, const cx<S> *__restrict__ p,
                                                            I i) {
  constexpr int VW = Sto<S>::VW;
  const cx<S> *__restrict__ q = p + bidx<S, I>(i, 0);
#pragma unroll
  for (int g = 0; g < 12 / VW; g++) load_group<S>(f + g * VW, q, g);
  if constexpr (Sto<S>::H) return *reinterpret_cast<const float *>(p + sidx<S, I>(i));
  else return (typename Sto<S>::C)1;
}

// acc += beta * (entry i of p), group by group (FP16 with the site's scale folded into beta)
template <typename S, typename I>
__device__ __forceinline__ void axpy_site(cx<typename Sto<S>::C> *acc, typename Sto<S>::C beta,
                                          const cx<S> *__restrict__ p, I i) {
  constexpr int VW = Sto<S>::VW;
  typedef typename Sto<S>::C T;
  const cx<S> *__restrict__ q = p + bidx<S, I>(i, 0);
  const T b = beta * site_scale<S, I>(p, i);
#pragma unroll
  for (int g = 0; g < 12 / VW; g++) {
    cx<T> v[VW];
    load_group<S>(v, q, g);
#pragma unroll
    for (int c = 0; c < VW; c++) {
      acc[g * VW + c].r = fma(b, v[c].r, acc[g * VW + c].r);
      acc[g * VW + c].i = fma(b, v[c].i, acc[g * VW + c].i);
    }
  }
}

// store the 12 components of entry i (FP16: with the site's scale); nt = non-temporal stores
template <typename S, typename I>
__device__ __forceinline__ void store_site(cx<S> *__restrict__ p, const cx<typename Sto<S>::C> *a, I i, bool nt) {
  constexpr int VW = Sto<S>::VW;
  cx<S> *__restrict__ q = p + bidx<S, I>(i, 0);
  if constexpr (Sto<S>::H) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 12; k++) m = fmaxf(m, fmaxf(fabsf(a[k].r), fabsf(a[k].i)));
    // a subnormal site maximum would overflow 1/m: scale the site by 2^64 (exact) before it
    const bool tiny = __builtin_expect(m > 0.f && m < FLT_MIN, 0);
    const float up = tiny ? 0x1p64f : 1.f;
    const float inv = m > 0.f ? 1.f / (m * up) : 0.f;
#pragma unroll
    for (int g = 0; g < 3; g++) {
      f16x8 v;
#pragma unroll
      for (int c = 0; c < 4; c++) {
        v[2 * c] = (half_t)((a[4 * g + c].r * up) * inv);
        v[2 * c + 1] = (half_t)((a[4 * g + c].i * up) * inv);
      }
      *reinterpret_cast<f16x8 *>(q + g * kBlk * VW) = v;
    }
    *reinterpret_cast<float *>(p + sidx<S, I>(i)) = m;
  } else if constexpr (VW == 1) {
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const f64x2 v = {a[k].r, a[k].i};
      f64x2 *d = reinterpret_cast<f64x2 *>(q + k * kBlk);
      if (nt) __builtin_nontemporal_store(v, d);
      else *d = v;
    }
  } else {
#pragma unroll
    for (int g = 0; g < 6; g++) {
      const f32x4 v = {a[2 * g].r, a[2 * g].i, a[2 * g + 1].r, a[2 * g + 1].i};
      f32x4 *d = reinterpret_cast<f32x4 *>(q + g * kBlk * VW);
      if (nt) __builtin_nontemporal_store(v, d);
      else *d = v;
    }
  }
}

// n / d for 0 <= n < 2^31 by multiply and shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994): l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1,
// n / d = (umulhi(n, m) + n) >> l.  The hop's site decomposition needs several divisions by lattice
// extents per lane; the hardware has no integer divide, so each would cost ~35 instructions.
struct FDiv {
  uint32_t d, m;
  int l;
};

static FDiv make_fdiv(uint32_t d) {
  FDiv f;
  f.d = d;
  f.l = 0;
  while ((1u << f.l) < d) f.l++;
  f.m = (uint32_t)((((uint64_t)1 << 32) * (((uint64_t)1 << f.l) - d)) / d + 1);
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FDiv &f) { return (__umulhi(n, f.m) + n) >> f.l; }

struct KGeom {
  int L[4];
  int Lxh;
  int64_t H4;
  int64_t stride;          // entries per checkerboard including block padding (Geom::half_pad)
  int64_t sY, sZ, sT;      // half-site strides of y, z, t
  // blocked sweep (tile map 4): blocks of Lxh x by x bz sites in one t-slice; a block's t-slices
  // are consecutive in the sweep (t fastest), blocks ordered y-block first
  int by, bz, nby;
  FDiv fLxh, fL1, fL2, fL3, fby, fnby, fbs;  // division by Lxh, Ly, Lz, Lt, by, nby, Lxh*by*bz
};

static void set_sweep_blocks(KGeom &k, int by, int bz) {
  k.by = by;
  k.bz = bz;
  k.nby = k.L[1] / by;
  k.fby = make_fdiv(by);
  k.fnby = make_fdiv(k.nby);
  k.fbs = make_fdiv((uint32_t)k.Lxh * by * bz);
}

static KGeom kgeom(const Geom &g) {
  KGeom k;
  for (int d = 0; d < 4; d++) k.L[d] = g.L[d];
  k.Lxh = g.L[0] / 2;
  k.H4 = g.H4;
  k.stride = g.half_pad();
  k.sY = k.Lxh;
  k.sZ = k.sY * g.L[1];
  k.sT = k.sZ * g.L[2];
  k.fLxh = make_fdiv(k.Lxh);
  k.fL1 = make_fdiv(g.L[1]);
  k.fL2 = make_fdiv(g.L[2]);
  k.fL3 = make_fdiv(g.L[3]);
  set_sweep_blocks(k, g.L[1], g.L[2]);
  return k;
}

// Site (half-lattice index h) visited at position r of the blocked sweep.
__device__ __forceinline__ uint32_t sweep_site(const KGeom &g, uint32_t r) {
  const uint32_t q = fdiv(r, g.fbs), w = r - q * g.fbs.d;   // block-slice, position in it
  const uint32_t yzb = fdiv(q, g.fL3), t = q - yzb * (uint32_t)g.L[3];
  const uint32_t bzi = fdiv(yzb, g.fnby), byi = yzb - bzi * (uint32_t)g.nby;
  const uint32_t w2 = fdiv(w, g.fLxh), xh = w - w2 * (uint32_t)g.Lxh;
  const uint32_t zz = fdiv(w2, g.fby), yy = w2 - zz * (uint32_t)g.by;
  const uint32_t y = byi * g.by + yy, z = bzi * g.bz + zz;
  return xh + (uint32_t)g.Lxh * (y + (uint32_t)g.L[1] * (z + (uint32_t)g.L[2] * t));
}

__device__ __forceinline__ bool gated(const Gate *gate) {
  return gate && (gate->k > gate->k_conv || gate->k > gate->kmax);
}

// x + sc h, x - sc h, x + i sc h, x - i sc h
template <typename T>
__device__ __forceinline__ cx<T> cadd_s(cx<T> x, cx<T> h, T sc) { return {fma(sc, h.r, x.r), fma(sc, h.i, x.i)}; }
template <typename T>
__device__ __forceinline__ cx<T> csub_s(cx<T> x, cx<T> h, T sc) { return {fma(-sc, h.r, x.r), fma(-sc, h.i, x.i)}; }
template <typename T>
__device__ __forceinline__ cx<T> caddi_s(cx<T> x, cx<T> h, T sc) { return {fma(-sc, h.i, x.r), fma(sc, h.r, x.i)}; }
template <typename T>
__device__ __forceinline__ cx<T> csubi_s(cx<T> x, cx<T> h, T sc) { return {fma(sc, h.i, x.r), fma(-sc, h.r, x.i)}; }


// spin projection (1 + gamma_mu) (PLUS) or (1 - gamma_mu) on a 12-component spinor f[spin*3+c],
// giving the 2-spinor h[spin*3+c], spin in {0,1}.  TwoSpinor.h:75-133
// MIX: every sum a +- b is formed as fma(+-one, b, a) with a unit `one` the compiler cannot see
// (the same rounded sum), so that when a and b are FP16 values widened to FP32 the widening folds
// into a mixed-precision FMA instead of two conversions and an add.
template <typename T, int MU, bool PLUS, bool MIX = false>
__device__ __forceinline__ void project(const cx<T> *f, cx<T> *h, T one = (T)1) {
#define GX_PADD(x, y) (MIX ? cadd_s<T>(x, y, one) : cadd<T>(x, y))
#define GX_PSUB(x, y) (MIX ? csub_s<T>(x, y, one) : csub<T>(x, y))
#define GX_PADDI(x, y) (MIX ? caddi_s<T>(x, y, one) : caddi<T>(x, y))
#define GX_PSUBI(x, y) (MIX ? csubi_s<T>(x, y, one) : csubi<T>(x, y))
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const cx<T> f0 = f[c], f1 = f[3 + c], f2 = f[6 + c], f3 = f[9 + c];
    if (MU == 0) {
      if (PLUS) { h[c] = GX_PADDI(f0, f3); h[3 + c] = GX_PADDI(f1, f2); }
      else      { h[c] = GX_PSUBI(f0, f3); h[3 + c] = GX_PSUBI(f1, f2); }
    } else if (MU == 1) {
      if (PLUS) { h[c] = GX_PSUB(f0, f3); h[3 + c] = GX_PADD(f1, f2); }
      else      { h[c] = GX_PADD(f0, f3); h[3 + c] = GX_PSUB(f1, f2); }
    } else if (MU == 2) {
      if (PLUS) { h[c] = GX_PADDI(f0, f2); h[3 + c] = GX_PSUBI(f1, f3); }
      else      { h[c] = GX_PSUBI(f0, f2); h[3 + c] = GX_PADDI(f1, f3); }
    } else {
      if (PLUS) { h[c] = GX_PADD(f0, f2); h[3 + c] = GX_PADD(f1, f3); }
      else      { h[c] = GX_PSUB(f0, f2); h[3 + c] = GX_PSUB(f1, f3); }
    }
  }
#undef GX_PADD
#undef GX_PSUB
#undef GX_PADDI
#undef GX_PSUBI
}

// accumulate the reconstruction of colour c of a projected 2-spinor whose spin rows are h0, h1,
// TwoSpinor.h:193-323
template <typename T, int MU, bool PLUS>
__device__ __forceinline__ void recon_colour(cx<T> *a, int c, cx<T> h0, cx<T> h1) {
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

// recon_colour of (sc h0, sc h1): the factor rides in the accumulation (one FMA per component
// instead of an add); FP16 storage, whose site scale is applied here
template <typename T, int MU, bool PLUS>
__device__ __forceinline__ void recon_colour_s(cx<T> *a, int c, cx<T> h0, cx<T> h1, T sc) {
#define add(x, h) cadd_s<T>(x, h, sc)
#define sub(x, h) csub_s<T>(x, h, sc)
#define addi(x, h) caddi_s<T>(x, h, sc)
#define subi(x, h) csubi_s<T>(x, h, sc)
  a[c] = add(a[c], h0);
  a[3 + c] = add(a[3 + c], h1);
  if (MU == 0) {
    if (PLUS) { a[6 + c] = subi(a[6 + c], h1); a[9 + c] = subi(a[9 + c], h0); }
    else      { a[6 + c] = addi(a[6 + c], h1); a[9 + c] = addi(a[9 + c], h0); }
  } else if (MU == 1) {
    if (PLUS) { a[6 + c] = add(a[6 + c], h1); a[9 + c] = sub(a[9 + c], h0); }
    else      { a[6 + c] = sub(a[6 + c], h1); a[9 + c] = add(a[9 + c], h0); }
  } else if (MU == 2) {
    if (PLUS) { a[6 + c] = subi(a[6 + c], h0); a[9 + c] = addi(a[9 + c], h1); }
    else      { a[6 + c] = addi(a[6 + c], h0); a[9 + c] = subi(a[9 + c], h1); }
  } else {
    if (PLUS) { a[6 + c] = add(a[6 + c], h0); a[9 + c] = add(a[9 + c], h1); }
    else      { a[6 + c] = sub(a[6 + c], h0); a[9 + c] = sub(a[9 + c], h1); }
  }
#undef add
#undef sub
#undef addi
#undef subi
}

// row r of U times the colour vector h (one spin row of the projected half-spinor)
template <typename T>
__device__ __forceinline__ cx<T> su3_row(const cx<T> *u, const cx<T> *h) {
  T re = 0, im = 0;
#pragma unroll
  for (int b = 0; b < 3; b++) {
    re = fma(u[b].r, h[b].r, re);
    re = fma(-u[b].i, h[b].i, re);
    im = fma(u[b].r, h[b].i, im);
    im = fma(u[b].i, h[b].r, im);
  }
  return {re, im};
}

// Links are staged per workgroup in LDS (one coalesced load, row stride kLinkPad complex per
// site to spread LDS banks) when the tile fits; otherwise each lane reads them from global
// memory (SoA over sites, coalesced across lanes when one lane = one site).
constexpr int kLinkPad = 73;

// Workgroup tile -> lattice mapping.  The hardware deals consecutive workgroups round-robin to the
// 8 XCDs.  map = 0: block b covers tile b (the XCDs interleave tile by tile).  map = 1: tile
// (b % 8) * per + b / 8 gives each XCD (and its L2) one contiguous slab of the lattice.  map = 2, 3:
// within every group of 8*R consecutive tiles (R = 8, 32) each XCD takes a run of R tiles, so the
// XCDs sweep the lattice together (shared Infinity Cache window) while each keeps runs in its L2.
// map = 4: tiles are runs of the blocked sweep (sweep_site) and the XCDs take slabs as in map 1.
// The grid is padded to a multiple of 8 (map 1, 4) or 8*R (map 2, 3) blocks; tiles past the end exit.
__device__ __forceinline__ uint32_t tile_of_block(uint32_t b, uint32_t nblocks, int map) {
  if (map == 1) {
    const uint32_t per = (nblocks + 7) >> 3;
    return (b & 7) * per + (b >> 3);
  }
  if (map == 2 || map == 3) {
    const uint32_t R = map == 2 ? 8 : 32;
    const uint32_t j = b % (8 * R);
    return (b - j) + (j & 7) * R + (j >> 3);
  }
  return b;
}

// Half-lattice indices of the 8 neighbours (x+, y+, z+, t+, x-, y-, z-, t-) of half-site h of
// checkerboard cb_out, with periodic wrap computed in place (no stencil table).
__device__ __forceinline__ void site_neighbours(const KGeom &g, uint32_t h, int cb_out, uint32_t *nb) {
  const uint32_t r1 = fdiv(h, g.fLxh), r2 = fdiv(r1, g.fL1), r3 = fdiv(r2, g.fL2);
  const int xh = (int)(h - r1 * (uint32_t)g.Lxh);
  const int y = (int)(r1 - r2 * (uint32_t)g.L[1]);
  const int z = (int)(r2 - r3 * (uint32_t)g.L[2]);
  const int t = (int)r3;
  const int x = 2 * xh + ((cb_out + y + z + t) & 1);
  const uint32_t base = h - xh;
  const int xp = (x + 1 == g.L[0]) ? 0 : x + 1;
  const int xm = (x == 0) ? g.L[0] - 1 : x - 1;
  const uint32_t sY = (uint32_t)g.sY, sZ = (uint32_t)g.sZ, sT = (uint32_t)g.sT;
  nb[0] = base + (xp >> 1);
  nb[4] = base + (xm >> 1);
  nb[1] = (y + 1 == g.L[1]) ? h - (uint32_t)(g.L[1] - 1) * sY : h + sY;
  nb[5] = (y == 0) ? h + (uint32_t)(g.L[1] - 1) * sY : h - sY;
  nb[2] = (z + 1 == g.L[2]) ? h - (uint32_t)(g.L[2] - 1) * sZ : h + sZ;
  nb[6] = (z == 0) ? h + (uint32_t)(g.L[2] - 1) * sZ : h - sZ;
  nb[3] = (t + 1 == g.L[3]) ? h - (uint32_t)(g.L[3] - 1) * sT : h + sT;
  nb[7] = (t == 0) ? h + (uint32_t)(g.L[3] - 1) * sT : h - sT;
}

// One leg on its neighbour spinor f: project to the half-spinor, multiply by the link and
// reconstruct, one colour row at a time (three link entries and two results live at once).
template <typename T, typename S, int MU, bool PLUS, bool LDSU>
__device__ __forceinline__ void leg_math(cx<T> *acc, const cx<T> *f, const cx<T> *ul, const cx<S> *__restrict__ U,
                                         int P, uint32_t h, uint32_t H4, T sc) {
  cx<T> hs[6];
  if constexpr (Sto<S>::H) {
    T one = (T)1;
    asm("" : "+s"(one));  // opaque to the optimiser: the projection's sums stay FMAs (see project)
    project<T, MU, PLUS, true>(f, hs, one);
  } else {
    project<T, MU, PLUS>(f, hs);
  }
#pragma unroll
  for (int r = 0; r < 3; r++) {
    cx<T> u[3];
#pragma unroll
    for (int b = 0; b < 3; b++) u[b] = LDSU ? ul[P * 9 + r * 3 + b] : ldc<S>(U, (P * 9 + r * 3 + b) * H4 + h);
    if constexpr (Sto<S>::H)  // f as stored; its site scale sc enters the accumulation
      recon_colour_s<T, MU, PLUS>(acc, r, su3_row<T>(u, hs), su3_row<T>(u, hs + 3), sc);
    else
      recon_colour<T, MU, PLUS>(acc, r, su3_row<T>(u, hs), su3_row<T>(u, hs + 3));
  }
}

// One leg: LEG-th entry of the leg sequence.  ORD = 0 is Grid's order (DaggerNo: Xm,Ym,Zm,Tm with
// (1+gamma) then Xp..Tp with (1-gamma); DaggerYes: Xp..Tp with (1+gamma) then Xm..Tm with
// (1-gamma)); ORD = 1 pairs the two legs of each direction so x-neighbour loads reuse L1.
template <bool DAG, int ORD, int LEG>
struct LegSel {
  static constexpr int MU = ORD == 0 ? (LEG & 3) : (LEG >> 1);
  static constexpr bool PLUS = ORD == 0 ? (LEG < 4) : ((LEG & 1) == 0);
  static constexpr int P = (PLUS != DAG) ? MU + 4 : MU;   // (1+gamma) legs read x-mu when !DAG
};

template <typename T, typename S, bool DAG, int ORD, int LEG, bool LDSU>
__device__ __forceinline__ void do_leg(cx<T> *acc, const cx<T> *f, const cx<T> *ul, const cx<S> *__restrict__ U,
                                       uint32_t h, uint32_t H4, T sc) {
  typedef LegSel<DAG, ORD, LEG> L;
  leg_math<T, S, L::MU, L::PLUS, LDSU>(acc, f, ul, U, L::P, h, H4, sc);
}

template <bool DAG, int ORD, int LEG>
__device__ __forceinline__ uint32_t leg_idx(const uint32_t *nb, uint32_t Ls, uint32_t s) {
  return nb[LegSel<DAG, ORD, LEG>::P] * Ls + s;
}

// The 8 legs of one output site.  The caller has issued the neighbour loads of leg 0 (into fa,
// its factor in sa) and, with PF, of leg 1 (into fb, sb); with PF the loads of leg L+1 are in
// flight while leg L is computed (sched_barrier keeps the compiler from sinking them).  Inputs
// stay as stored (load_site_raw): FP16 site scales enter the reconstruction.
template <typename T, typename S, bool DAG, bool PF, int ORD, bool LDSU>
__device__ __forceinline__ void hop_site(cx<T> *acc, cx<T> *fa, cx<T> *fb, T &sa, T &sb, const cx<S> *__restrict__ in,
                                         const cx<S> *__restrict__ U, const cx<T> *ul, const uint32_t *nb,
                                         uint32_t LS, uint32_t s, uint32_t h, uint32_t H4) {
#define IDX(L) leg_idx<DAG, ORD, L>(nb, LS, s)
#define LEG(L, F, SC) do_leg<T, S, DAG, ORD, L, LDSU>(acc, F, ul, U, h, H4, SC)
  if (PF) {
    __builtin_amdgcn_sched_barrier(0);
    LEG(0, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(2));
    __builtin_amdgcn_sched_barrier(0);
    LEG(1, fb, sb);
    sb = load_site_raw<S>(fb, in, IDX(3));
    __builtin_amdgcn_sched_barrier(0);
    LEG(2, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(4));
    __builtin_amdgcn_sched_barrier(0);
    LEG(3, fb, sb);
    sb = load_site_raw<S>(fb, in, IDX(5));
    __builtin_amdgcn_sched_barrier(0);
    LEG(4, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(6));
    __builtin_amdgcn_sched_barrier(0);
    LEG(5, fb, sb);
    sb = load_site_raw<S>(fb, in, IDX(7));
    __builtin_amdgcn_sched_barrier(0);
    LEG(6, fa, sa);
    LEG(7, fb, sb);
  } else {
    LEG(0, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(1));
    LEG(1, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(2));
    LEG(2, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(3));
    LEG(3, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(4));
    LEG(4, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(5));
    LEG(5, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(6));
    LEG(6, fa, sa);
    sa = load_site_raw<S>(fa, in, IDX(7));
    LEG(7, fa, sa);
  }
#undef IDX
#undef LEG
}

// Epilogue of the hop.  EPI 0: out = hop; 1: out_s = alpha_s hop_s + beta_s phi_s (per-slice
// coefficients, so one multi-RHS launch can carry a different mass per right-hand side);
// 2: out = alpha A(hop); 3: out = alpha A(hop) + beta phi, with A a per-chirality Ls x Ls matrix applied across the s-row
// of each site from LDS and phi in the layout of out.  Whether phi is read is a template choice: a
// run-time null test inside the unrolled epilogue makes the compiler spill the whole hop.
template <typename S>
struct HopEpi {
  typedef typename Sto<S>::C T;
  const cx<S> *phi;
  const T *mat;   // EPI 2, 3: [2][LS][LS], chirality 0 = spins 0,1
  T alpha, beta;  // EPI 2, 3
  const T *coef;  // EPI 1, 5, 6: per-slice coefficients, alpha_s = coef[s], beta_s = coef[LS + s]
  // k_dhop only.  EPI 5: out as EPI 1, and the per-tile slice sums of |out|^2 in npart [tile][LS];
  // EPI 6 (EPI 1 result) and 7 (EPI 3 result): rv_s -= acoef[astride s] (result)_s in place of
  // storing out, per-tile slice sums of |rv_s|^2
  double *npart;
  cx<S> *rv;
  const double *acoef;
  int astride;
};

// One lane per (4D half-site h, s): s fastest, 64/LS sites per wave, 4 waves per workgroup, so a
// wave reads each neighbour's s-row contiguously and the lanes of one site share its links.
// flags: bit 0 = non-temporal output stores.  ORD: leg order (0 Grid's, 1 paired directions).
// Waves per SIMD the compiler must leave VGPRs for (512 / waves each).  The FP16-storage hop with
// prefetch in Grid's leg order fits three waves (168 VGPRs) with at most two spilled against 174
// unbounded (two waves), which made its 12-RHS form 16% and the 16^4 x 16 one 8% faster; the other
// FP16 forms (and Ls < 4) spill up to 35 VGPRs under the same bound and stay unbounded.
template <typename S, int LS, bool PF, int ORD>
constexpr int dhop_waves() {
  return (Sto<S>::H && PF && ORD == 0 && LS >= 4) ? 3 : 1;
}

template <typename S, int LS, bool DAG, bool PF, int ORD, int EPI>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(dhop_waves<S, LS, PF, ORD>())))
k_dhop(const cx<S> *__restrict__ in, cx<S> *__restrict__ out,
                                              const cx<S> *__restrict__ U, KGeom g, int cb_out,
                                              const Gate *__restrict__ gate, uint32_t ntiles, int map, int flags,
                                              HopEpi<S> ep) {
  typedef typename Sto<S>::C T;
  if (gated(gate)) return;
  constexpr int SPW = 64 / LS;
  constexpr int SPB = 4 * SPW;
  constexpr bool LDSU = (LS > 1) && (SPB * kLinkPad * sizeof(cx<T>) <= 40960);
  constexpr int NLINK = LDSU ? SPB * kLinkPad : 1;
  constexpr bool MAT = EPI == 2 || EPI == 3 || EPI == 7, PHI = EPI == 1 || EPI == 3 || EPI >= 5;
  constexpr bool NORM = EPI >= 5, RUPD = EPI == 6 || EPI == 7;
  constexpr int NSTAGE = MAT ? 4 * 12 * 64 : 1;
  __shared__ cx<T> lds[NLINK > NSTAGE ? NLINK : NSTAGE];  // links, then (MAT) the hop results
  const bool sweep = map == 4;  // blocked sweep: tiles are runs of the sweep order, XCD slabs
  const uint32_t tile = tile_of_block(blockIdx.x, gridDim.x, sweep ? 1 : map);
  if (tile >= ntiles) return;  // uniform across the workgroup
  const uint32_t H4 = (uint32_t)g.H4;
  const uint32_t h0 = tile * SPB;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int ls = lane / LS;
  const int s = lane - ls * LS;
  const int lsb = w * SPW + ls;
  const bool valid = (ls < SPW) && (h0 + lsb < H4);
  // idle lanes compute in-bounds addresses and skip the work
  const uint32_t h = sweep ? sweep_site(g, valid ? h0 + lsb : h0) : (valid ? h0 + lsb : h0);
  uint32_t nb[8];
  site_neighbours(g, h, cb_out, nb);
  // Link tile of the workgroup: global -> registers first, then the first neighbour loads, then
  // registers -> LDS and a barrier that waits for LDS traffic only, so the neighbour loads stay in
  // flight across it.
  // Thread t < G * SPB stages link entries pk = t / SPB + G j of tile site t % SPB (G = 256 / SPB
  // groups), so it finds its site's lattice index once; the lanes of a group read consecutive sites.
  constexpr int G = 256 / SPB;
  constexpr int NU = LDSU ? (72 + G - 1) / G : 1;
  const uint32_t ugrp = threadIdx.x / SPB, usite = threadIdx.x - ugrp * SPB;
  const bool ustage = LDSU && ugrp < (uint32_t)G && h0 + usite < H4;
  cx<T> ureg[NU];
  if (LDSU) {
    const uint32_t hs = sweep && ustage ? sweep_site(g, h0 + usite) : h0 + usite;
#pragma unroll
    for (int j = 0; j < NU; j++) {
      const uint32_t pk = ugrp + G * j;
      ureg[j] = ustage && pk < 72 ? ldc<S>(U, pk * H4 + hs) : cx<T>{0, 0};
    }
  }
  cx<T> fa[12], fb[12];
  T sa = load_site_raw<S>(fa, in, leg_idx<DAG, ORD, 0>(nb, LS, s)), sb = (T)1;
  if (PF) sb = load_site_raw<S>(fb, in, leg_idx<DAG, ORD, 1>(nb, LS, s));
  if (LDSU) {
    if (ugrp < (uint32_t)G) {
#pragma unroll
      for (int j = 0; j < NU; j++) {
        const uint32_t pk = ugrp + G * j;
        if (pk < 72) lds[usite * kLinkPad + pk] = ureg[j];
      }
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  }
  if (!MAT && !NORM && !valid) return;
  const cx<T> *ul = lds + (LDSU ? lsb * kLinkPad : 0);
  cx<T> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0, 0};
  if (valid) hop_site<T, S, DAG, PF, ORD, LDSU>(acc, fa, fb, sa, sb, in, U, ul, nb, LS, s, h, H4);
  const uint32_t o = h * LS + s;
  if (MAT) {
    // every wave has finished reading links: reuse the LDS for the hop results [w][k][lane]
    __syncthreads();
    cx<T> *stg = lds + w * 12 * 64;
#pragma unroll
    for (int k = 0; k < 12; k++) stg[k * 64 + lane] = acc[k];
    __syncthreads();
    if (!valid && !NORM) return;
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = {0, 0};
    if (valid) {
      const T *ap = ep.mat + s * LS, *am = ep.mat + LS * LS + s * LS;
#pragma unroll
      for (int tt = 0; tt < LS; tt++) {
        const T cp = ap[tt], cm = am[tt];
#pragma unroll
        for (int k = 0; k < 12; k++) {
          const cx<T> v = stg[k * 64 + ls * LS + tt];
          const T c = k < 6 ? cp : cm;
          acc[k].r = fma(c, v.r, acc[k].r);
          acc[k].i = fma(c, v.i, acc[k].i);
        }
      }
    }
  }
  if (EPI != 0 && valid) {
    const bool pc = EPI == 1 || EPI == 5 || EPI == 6;
    const T al = pc ? ep.coef[s] : ep.alpha, be = pc ? ep.coef[LS + s] : ep.beta;
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = {al * acc[k].r, al * acc[k].i};
    if (PHI) axpy_site<S>(acc, be, ep.phi, o);
  }
  if constexpr (NORM) {
    // per-slice |out|^2 (EPI 5) or |rv|^2 (EPI 6) of the tile's sites: lane sums in FP64, then the
    // tile's sites in fixed order
    __shared__ double nsh[256];
    double nv = 0;
    if (valid) {
      if (RUPD) {
        cx<T> rv[12];
        load_site<S>(rv, ep.rv, o);
        const T a = (T)(-ep.acoef[ep.astride * s]);
#pragma unroll
        for (int k = 0; k < 12; k++) {
          rv[k] = {fma(a, acc[k].r, rv[k].r), fma(a, acc[k].i, rv[k].i)};
          nv += (double)fma(rv[k].r, rv[k].r, rv[k].i * rv[k].i);
        }
        store_site<S>(ep.rv, rv, o, false);
      } else {
#pragma unroll
        for (int k = 0; k < 12; k++) nv += (double)fma(acc[k].r, acc[k].r, acc[k].i * acc[k].i);
        store_site<S>(out, acc, o, (flags & 1) != 0);
      }
    }
    nsh[threadIdx.x] = nv;
    __syncthreads();
    if (threadIdx.x < LS) {
      double v = 0;
      for (int ww = 0; ww < 4; ww++)
        for (int j = 0; j < SPW; j++) v += nsh[ww * 64 + j * LS + threadIdx.x];
      ep.npart[(size_t)tile * LS + threadIdx.x] = v;
    }
    return;
  }
  store_site<S>(out, acc, o, (flags & 1) != 0);
}

// ---------------- general fifth-dimension extent ----------------
// Tile shape of the general kernels: a workgroup of W waves covers EPB = 64 W consecutive positions
// of the site order times LS slices, i.e. SPB = EPB / LS whole sites, and lane t of the workgroup
// works on slice t % LS of tile site t / LS.  With the linear tile maps every wave reads whole
// aligned 64-entry storage blocks and no lane idles, for any LS with W a multiple of
// LS / gcd(LS, 64).  LS, W and the group LB of the fifth-dimension matrices (block diagonal in
// groups of LB consecutive slices; LB = LS for one Mobius operator) are run-time values.  Links
// are staged per workgroup in LDS (LS >= 8: at most 9 link entries per thread).
struct GShape {
  uint32_t LS, LB, SPB;
  FDiv fLS, fLB;
};
constexpr int kGenMaxWaves = 8;
constexpr int kGenMinLs = 8;
constexpr int kGenLinkPerThread = (72 + kGenMinLs - 1) / kGenMinLs;

// the matrix epilogue across the s-rows of the workgroup's sites, one chirality at a time: the 6
// components of chirality chi of acc are staged in LDS (6 columns of EPB), then acc[6 chi + k] =
// sum over the slices tt of this lane's group of A_chi[s_g][tt] * staged[k][site, group + tt]
template <typename T>
__device__ __forceinline__ void group_matrix(cx<T> *acc, cx<T> *stg, const T *mat, const GShape &sh, uint32_t t,
                                             uint32_t hl, uint32_t s, uint32_t epb) {
  const uint32_t sg = s - fdiv(s, sh.fLB) * sh.LB;   // slice within the group
  const uint32_t row = hl * sh.LS + (s - sg);        // first slice of the group in the tile
#pragma unroll
  for (int chi = 0; chi < 2; chi++) {
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; k++) stg[k * epb + t] = acc[6 * chi + k];
    __syncthreads();
    cx<T> r[6];
#pragma unroll
    for (int k = 0; k < 6; k++) r[k] = {0, 0};
    const T *a = mat + (size_t)chi * sh.LB * sh.LB + (size_t)sg * sh.LB;
    for (uint32_t tt = 0; tt < sh.LB; tt++) {
      const T c = a[tt];
#pragma unroll
      for (int k = 0; k < 6; k++) {
        const cx<T> v = stg[k * epb + row + tt];
        r[k].r = fma(c, v.r, r[k].r);
        r[k].i = fma(c, v.i, r[k].i);
      }
    }
#pragma unroll
    for (int k = 0; k < 6; k++) acc[6 * chi + k] = r[k];
  }
}

// The hop of k_dhop (same legs, projectors, link staging and epilogues) in the general tile shape,
// for one tile.  Every thread of the workgroup calls it (it synchronises the workgroup).
template <typename S, bool DAG, int EPI>
__device__ __forceinline__ void hop_tile(const cx<S> *__restrict__ in, cx<S> *__restrict__ out,
                                         const cx<S> *__restrict__ U, const KGeom &g, int cb_out, uint32_t tile,
                                         bool sweep, int flags, const HopEpi<S> &ep, const GShape &sh,
                                         cx<typename Sto<S>::C> *lds) {
  typedef typename Sto<S>::C T;
  constexpr bool MAT = EPI == 2 || EPI == 3 || EPI == 7, PHI = EPI == 1 || EPI == 3 || EPI >= 5;
  constexpr bool NORM = EPI >= 5, RUPD = EPI == 6 || EPI == 7;
  const uint32_t H4 = (uint32_t)g.H4, LS = sh.LS, epb = blockDim.x;
  const uint32_t t = threadIdx.x;
  const uint32_t hl = fdiv(t, sh.fLS), s = t - hl * LS;
  const uint32_t p0 = tile * sh.SPB;
  const bool valid = p0 + hl < H4;
  // lanes past the lattice take the tile's first site: in-bounds addresses, work skipped
  const uint32_t pos = valid ? p0 + hl : p0;
  const uint32_t h = sweep ? sweep_site(g, pos) : pos;
  uint32_t nb[8];
  site_neighbours(g, h, cb_out, nb);
  // links of tile site hl: thread (hl, s) loads entries s, s + LS, ... (< 72)
  cx<T> ureg[kGenLinkPerThread];
#pragma unroll
  for (int j = 0; j < kGenLinkPerThread; j++) {
    const uint32_t pk = s + LS * j;
    ureg[j] = pk < 72 ? ldc<S>(U, pk * H4 + h) : cx<T>{0, 0};
  }
  cx<T> fa[12], fb[12];
  T sa = load_site_raw<S>(fa, in, leg_idx<DAG, 0, 0>(nb, LS, s));
  T sb = load_site_raw<S>(fb, in, leg_idx<DAG, 0, 1>(nb, LS, s));
#pragma unroll
  for (int j = 0; j < kGenLinkPerThread; j++) {
    const uint32_t pk = s + LS * j;
    if (pk < 72) lds[hl * kLinkPad + pk] = ureg[j];
  }
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
  cx<T> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0, 0};
  if (valid) hop_site<T, S, DAG, true, 0, true>(acc, fa, fb, sa, sb, in, U, lds + hl * kLinkPad, nb, LS, s, h, H4);
  if (MAT) group_matrix<T>(acc, lds, ep.mat, sh, t, hl, s, epb);
  double nv = 0;
  if (valid) {
    const uint32_t o = h * LS + s;
    if (EPI != 0) {
      const bool pc = EPI == 1 || EPI == 5 || EPI == 6;
      const T al = pc ? ep.coef[s] : ep.alpha, be = pc ? ep.coef[LS + s] : ep.beta;
#pragma unroll
      for (int k = 0; k < 12; k++) acc[k] = {al * acc[k].r, al * acc[k].i};
      if (PHI) axpy_site<S>(acc, be, ep.phi, o);
    }
    if (RUPD) {
      // the residual update in place of storing the result (k_dhop's EPI 6 / 7)
      cx<T> rv[12];
      load_site<S>(rv, ep.rv, o);
      const T a = (T)(-ep.acoef[ep.astride * s]);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        rv[k] = {fma(a, acc[k].r, rv[k].r), fma(a, acc[k].i, rv[k].i)};
        nv += (double)fma(rv[k].r, rv[k].r, rv[k].i * rv[k].i);
      }
      store_site<S>(ep.rv, rv, o, false);
    } else {
      if (NORM)
#pragma unroll
        for (int k = 0; k < 12; k++) nv += (double)fma(acc[k].r, acc[k].r, acc[k].i * acc[k].i);
      store_site<S>(out, acc, o, (flags & 1) != 0);
    }
  }
  if constexpr (NORM) {
    // per-slice sums over the tile's sites in fixed order (the persistent loop's barrier at the
    // next tile orders these reads before the next writes)
    __shared__ double nsh[64 * kGenMaxWaves];
    nsh[t] = nv;
    __syncthreads();
    if (t < LS) {
      double v = 0;
      for (uint32_t j = 0; j < sh.SPB; j++) v += nsh[j * LS + t];
      ep.npart[(size_t)tile * LS + t] = v;
    }
  }
}

// One workgroup per tile (persist = 0), or persist = 1: a grid of resident workgroups (a multiple
// of 8, so a workgroup keeps its XCD) each looping over the virtual blocks b, b + gridDim, ... of
// the same tile map, nblocks of them: the same tiles and arithmetic with fewer workgroup launches.
template <typename S, bool DAG, int EPI>
__global__ void __launch_bounds__(64 * kGenMaxWaves)
    k_dhop_g(const cx<S> *__restrict__ in, cx<S> *__restrict__ out, const cx<S> *__restrict__ U, KGeom g, int cb_out,
             const Gate *__restrict__ gate, uint32_t ntiles, int map, int flags, HopEpi<S> ep, GShape sh,
             uint32_t nblocks, int persist) {
  typedef typename Sto<S>::C T;
  if (gated(gate)) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(lds_raw);
  const bool sweep = map == 4;
  const int tmap = sweep ? 1 : map;
  if (!persist) {
    const uint32_t tile = tile_of_block(blockIdx.x, gridDim.x, tmap);
    if (tile < ntiles) hop_tile<S, DAG, EPI>(in, out, U, g, cb_out, tile, sweep, flags, ep, sh, lds);
    return;
  }
  for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const uint32_t tile = tile_of_block(b, nblocks, tmap);
    if (tile >= ntiles) continue;  // uniform across the workgroup
    __syncthreads();               // the previous tile's LDS reads are done
    hop_tile<S, DAG, EPI>(in, out, U, g, cb_out, tile, sweep, flags, ep, sh, lds);
  }
}

// ---------------- hop with x, y and z legs from a tile of input rows in LDS ----------------
// Tile = the 2 x 2 rows (y0, y0 + 1) x (z0, z0 + 1) of one t-slice, every x and every slice s:
// RW = Lxh LS entries per row and 4 RW threads, thread t on row r = t / RW (y = y0 + (r & 1),
// z = z0 + (r >> 1)) and entry w = t % RW of it (site xh = w / LS, slice s = w % LS).  Each thread
// first loads the input entry at its own position: the four input rows of the tile (opposite
// checkerboard) land in LDS, one 16-byte group after another.  The x legs of every lane, and the
// y and z legs whose neighbour row is in the tile, read their neighbour there; the t legs and the
// y / z legs at the tile's faces load from global memory.  That is 5 global neighbour loads per
// output entry instead of 8 (1 own + 1 y face + 1 z face + 2 t), for kernels bound by the vector
// memory pipeline (the address and data units of each CU) rather than by HBM.  Every lane of a row
// has the same sources, so each of the four rows runs its own compiled leg sequence.  Tiles are
// ordered t fastest (maps 0-3), or in XCD boxes swept t-slice by t-slice (map 4).  Same legs, leg
// order, link products and epilogues as k_dhop: the results are bitwise those of the other tile
// shapes.  Needs Ly, Lz >= 4.
struct RowTile {
  uint32_t LS, RW, nyb;   // slices, entries per row, y blocks (Ly / 2)
  FDiv fLS, fRW, fLt, fnyb, fQ;  // division by LS, RW, Lt, nyb, 4 Lxh (tile sites)
  // XCD blocks (xblk = 1): the tiles split into 8 boxes of bY x bZ tile columns x tper t-slices
  // (pY x pZ x pT boxes), one per XCD (per = ntiles / 8 tiles each, the map-1 slab of that XCD),
  // swept t-slice by t-slice with the box's columns innermost, so a tile's y, z and t neighbour
  // rows inside the box were loaded shortly before by the same XCD
  int xblk;
  uint32_t per, nc, bY, bZ, tper, pY, pZ;
  FDiv fper, fnc, fbY, fpY, fpZ;
};

template <typename T>
__device__ __forceinline__ void lds_load_site(cx<T> *f, const cx<T> *lin, uint32_t e, uint32_t nent) {
  constexpr int VWL = 16 / (int)sizeof(cx<T>);
  const cx<T> *q = lin + (size_t)e * VWL;
#pragma unroll
  for (int g = 0; g < 12 / VWL; g++) {
#pragma unroll
    for (int c = 0; c < VWL; c++) f[g * VWL + c] = q[(size_t)g * nent * VWL + c];
  }
}

template <typename T>
__device__ __forceinline__ void lds_store_site(cx<T> *lin, const cx<T> *f, uint32_t e, uint32_t nent) {
  constexpr int VWL = 16 / (int)sizeof(cx<T>);
  cx<T> *q = lin + (size_t)e * VWL;
#pragma unroll
  for (int g = 0; g < 12 / VWL; g++) {
#pragma unroll
    for (int c = 0; c < VWL; c++) q[(size_t)g * nent * VWL + c] = f[g * VWL + c];
  }
}

// up to 1024 threads (16 waves, 128 VGPRs) for FP32 arithmetic, 512 for FP64
template <typename S>
constexpr int rows_max_threads() { return sizeof(typename Sto<S>::C) == 8 ? 512 : 1024; }

// The 8 legs of one lane of a row tile whose row has y - y0 = RY and z - z0 = RZ (compile time, so
// every leg's source is fixed: x legs and the y / z legs towards the tile's other row from LDS, the
// t legs and the y / z legs across the tile's faces from global memory, one global leg ahead).
template <typename T, typename S, bool DAG, int RY, int RZ>
__device__ __forceinline__ void rows_legs(cx<T> *acc, const cx<S> *__restrict__ in, const cx<S> *__restrict__ U,
                                          const cx<T> *lin, const T *lsc, const cx<T> *ul, const uint32_t *nb,
                                          uint32_t rowh, uint32_t r, uint32_t w, uint32_t RW, uint32_t nent, uint32_t LS,
                                          uint32_t s, uint32_t h, uint32_t H4) {
  // neighbour P of leg L (Grid's order) and whether it is in the tile
  constexpr auto P = [](int L) constexpr { return L < 4 ? ((DAG ? 0 : 4) + (L & 3)) : ((DAG ? 4 : 0) + (L & 3)); };
  constexpr auto in_tile = [](int p) constexpr {
    return p == 0 || p == 4 || (p == 1 && RY == 0) || (p == 5 && RY == 1) || (p == 2 && RZ == 0) ||
           (p == 6 && RZ == 1);
  };
  // tile entry of an in-tile neighbour
  auto entry = [&](int p) -> uint32_t {
    if (p == 0 || p == 4) return r * RW + (nb[p] - rowh) * LS + s;
    if (p == 1 || p == 5) return (r ^ 1u) * RW + w;
    return (r ^ 2u) * RW + w;
  };
  cx<T> fg[12];
  // the first global leg's neighbour, loaded before the LDS legs that precede it
  constexpr int G0 = !in_tile(P(0)) ? 0 : !in_tile(P(1)) ? 1 : !in_tile(P(2)) ? 2 : !in_tile(P(3)) ? 3
                     : !in_tile(P(4)) ? 4 : !in_tile(P(5)) ? 5 : !in_tile(P(6)) ? 6 : 7;
  T sg = load_site_raw<S>(fg, in, nb[P(G0)] * LS + s);
#define GX_ROW_LEG(L)                                                                                  \
  {                                                                                                    \
    constexpr int PL = P(L);                                                                           \
    cx<T> f[12];                                                                                       \
    T sc;                                                                                              \
    if constexpr (in_tile(PL)) {                                                                       \
      const uint32_t en = entry(PL);                                                                   \
      lds_load_site<T>(f, lin, en, nent);                                                              \
      if constexpr (Sto<S>::H) sc = lsc[en];                                                           \
      else sc = (T)1;                                                                                  \
    } else {                                                                                           \
      _Pragma("unroll") for (int k = 0; k < 12; k++) f[k] = fg[k];                                     \
      sc = sg;                                                                                         \
      constexpr int NX = L >= 7 ? -1 : !in_tile(P(L + 1 > 7 ? 7 : L + 1)) ? L + 1                      \
                       : L + 2 <= 7 && !in_tile(P(L + 2 > 7 ? 7 : L + 2)) ? L + 2                      \
                       : L + 3 <= 7 && !in_tile(P(L + 3 > 7 ? 7 : L + 3)) ? L + 3                      \
                       : L + 4 <= 7 && !in_tile(P(L + 4 > 7 ? 7 : L + 4)) ? L + 4                      \
                       : L + 5 <= 7 && !in_tile(P(L + 5 > 7 ? 7 : L + 5)) ? L + 5                      \
                       : L + 6 <= 7 && !in_tile(P(L + 6 > 7 ? 7 : L + 6)) ? L + 6 : -1;                \
      if constexpr (NX >= 0) sg = load_site_raw<S>(fg, in, nb[P(NX)] * LS + s);                        \
    }                                                                                                  \
    do_leg<T, S, DAG, 0, L, true>(acc, f, ul, U, h, H4, sc);                                           \
  }
  GX_ROW_LEG(0) GX_ROW_LEG(1) GX_ROW_LEG(2) GX_ROW_LEG(3) GX_ROW_LEG(4) GX_ROW_LEG(5) GX_ROW_LEG(6) GX_ROW_LEG(7)
#undef GX_ROW_LEG
}

template <typename S, bool DAG, int EPI>
__global__ void __launch_bounds__(rows_max_threads<S>())
    k_dhop_rows(const cx<S> *__restrict__ in, cx<S> *__restrict__ out, const cx<S> *__restrict__ U, KGeom g,
                int cb_out, const Gate *__restrict__ gate, uint32_t ntiles, int map, int flags, HopEpi<S> ep,
                GShape sh, RowTile rt) {
  typedef typename Sto<S>::C T;
  constexpr bool MAT = EPI >= 2, PHI = (EPI & 1) != 0;
  if (gated(gate)) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t nent = 4 * rt.RW;                 // entries of the tile (= threads)
  cx<T> *lin = reinterpret_cast<cx<T> *>(lds_raw);  // [12 / VWL groups][nent][VWL]
  cx<T> *links = lin + (size_t)12 * nent;           // [4 Lxh sites][kLinkPad]
  T *lsc = reinterpret_cast<T *>(links + (size_t)4 * g.Lxh * kLinkPad);  // FP16: the entries' site scales
  const uint32_t tile = tile_of_block(blockIdx.x, gridDim.x, map);
  if (tile >= ntiles) return;  // uniform across the workgroup
  const uint32_t H4 = (uint32_t)g.H4, LS = rt.LS, Lxh = (uint32_t)g.Lxh;
  uint32_t yb, zb, tslice;
  if (rt.xblk) {
    const uint32_t xcd = fdiv(tile, rt.fper), l = tile - xcd * rt.per;
    const uint32_t tl = fdiv(l, rt.fnc), cl = l - tl * rt.nc;
    const uint32_t cz = fdiv(cl, rt.fbY), cy = cl - cz * rt.bY;
    const uint32_t q = fdiv(xcd, rt.fpY), iy = xcd - q * rt.pY;
    const uint32_t it = fdiv(q, rt.fpZ), iz = q - it * rt.pZ;
    yb = iy * rt.bY + cy;
    zb = iz * rt.bZ + cz;
    tslice = it * rt.tper + tl;
  } else {
    const uint32_t tt = fdiv(tile, rt.fLt);
    tslice = tile - tt * (uint32_t)g.L[3];
    zb = fdiv(tt, rt.fnyb);
    yb = tt - zb * rt.nyb;
  }
  const uint32_t y0 = 2 * yb, z0 = 2 * zb;
  const uint32_t t = threadIdx.x;
  const uint32_t r = fdiv(t, rt.fRW), w = t - r * rt.RW;
  const uint32_t xh = fdiv(w, rt.fLS), s = w - xh * LS;
  const uint32_t y = y0 + (r & 1), z = z0 + (r >> 1);
  const uint32_t rowh = Lxh * (y + (uint32_t)g.L[1] * (z + (uint32_t)g.L[2] * tslice));  // h of x = 0
  const uint32_t h = rowh + xh;
  // the tile's links: 72 entries of each of its 4 Lxh sites
  const uint32_t nsite = 4 * Lxh;
  for (uint32_t i = t; i < 72 * nsite; i += blockDim.x) {
    const uint32_t pk = fdiv(i, rt.fQ), q = i - pk * nsite;
    const uint32_t qr = fdiv(q, g.fLxh), qx = q - qr * Lxh;
    const uint32_t hq = qx + Lxh * (y0 + (qr & 1) + (uint32_t)g.L[1] * (z0 + (qr >> 1) + (uint32_t)g.L[2] * tslice));
    links[q * kLinkPad + pk] = ldc<S>(U, pk * H4 + hq);
  }
  // own input entry -> LDS (as stored, with its FP16 site scale beside it)
  {
    cx<T> f[12];
    const T sc = load_site_raw<S>(f, in, h * LS + s);
    lds_store_site<T>(lin, f, t, nent);
    if constexpr (Sto<S>::H) lsc[t] = sc;
  }
  uint32_t nb[8];
  site_neighbours(g, h, cb_out, nb);
  __syncthreads();
  const cx<T> *ul = links + (t / LS) * kLinkPad;
  cx<T> acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0, 0};
  switch (r) {  // uniform over a wave when a row is a multiple of 64 entries
    case 0: rows_legs<T, S, DAG, 0, 0>(acc, in, U, lin, lsc, ul, nb, rowh, r, w, rt.RW, nent, LS, s, h, H4); break;
    case 1: rows_legs<T, S, DAG, 1, 0>(acc, in, U, lin, lsc, ul, nb, rowh, r, w, rt.RW, nent, LS, s, h, H4); break;
    case 2: rows_legs<T, S, DAG, 0, 1>(acc, in, U, lin, lsc, ul, nb, rowh, r, w, rt.RW, nent, LS, s, h, H4); break;
    default: rows_legs<T, S, DAG, 1, 1>(acc, in, U, lin, lsc, ul, nb, rowh, r, w, rt.RW, nent, LS, s, h, H4); break;
  }
  if (MAT) group_matrix<T>(acc, lin, ep.mat, sh, t, t / LS, s, blockDim.x);
  const uint32_t o = h * LS + s;
  if (EPI != 0) {
    const T al = EPI == 1 ? ep.coef[s] : ep.alpha, be = EPI == 1 ? ep.coef[LS + s] : ep.beta;
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = {al * acc[k].r, al * acc[k].i};
    if (PHI) axpy_site<S>(acc, be, ep.phi, o);
  }
  store_site<S>(out, acc, o, (flags & 1) != 0);
}

// ---------------- multi-RHS Wilson hop with the link products on the matrix cores ----------------
// For a multi-RHS operator (Ls = Nrhs <= 16 right-hand sides sharing the links) the product of each
// leg is a small dense contraction: chi[spin][a][n] = sum_b U[a][b] h[spin][b][n] over colours
// a, b and right-hand sides n.  In real form it is D (6 x 16) = A (6 x 6) B (6 x 16), rows
// m = 2 a + re/im, columns k = 2 b + re/im, A the realified link [[Ur, -Ui], [Ui, Ur]] and B the
// projected half-spinors of the 16 right-hand sides, computed per (site, spin) with two
// v_mfma_f32_16x16x4f32 (K = 8, rows and columns past 6 zero).  Operand layout of that instruction
// (lane l of the wave): A[l % 16][l / 16], B[l / 16][l % 16], D[4 (l / 16) + v][l % 16] in result
// register v.  Each leg stages the lanes' half-spinors in LDS in B's layout, the waves run the
// (site, spin) products, and the lanes read their results back for the reconstruction.  FP32
// storage, general tile shape (whole sites per workgroup), Grid's leg order.
typedef float f32x4v __attribute__((ext_vector_type(4)));
constexpr int kMfmaRows = 8;   // K and M padded to 8 (6 used) in the LDS staging

template <bool DAG, int LEG>
__device__ __forceinline__ void mfma_leg(cx<float> *acc, const cx<float> *f, float *hsl, float *dl,
                                         const cx<float> *links, uint32_t hl, uint32_t s, uint32_t LS, uint32_t spb) {
  typedef LegSel<DAG, 0, LEG> L;
  cx<float> hs[6];
  project<float, L::MU, L::PLUS>(f, hs);
#pragma unroll
  for (int sp = 0; sp < 2; sp++)
#pragma unroll
    for (int c = 0; c < 3; c++) {
      float *row = hsl + ((hl * 2 + sp) * kMfmaRows + 2 * c) * 16 + s;
      row[0] = hs[sp * 3 + c].r;
      row[16] = hs[sp * 3 + c].i;
    }
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwaves = blockDim.x >> 6;
  const int i = lane & 15, q = lane >> 4;
  for (uint32_t job = wave; job < 2 * spb; job += nwaves) {
    const uint32_t site = job >> 1, sp = job & 1;
    const cx<float> *u = links + site * kLinkPad + L::P * 9;
    const float *B = hsl + (site * 2 + sp) * kMfmaRows * 16;
    f32x4v d = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int k = 4 * kk + q;
      float av = 0.f, bv = 0.f;
      if (i < 6 && k < 6) {
        const cx<float> uab = u[(i >> 1) * 3 + (k >> 1)];
        const int ra = i & 1, rb = k & 1;
        av = ra == rb ? uab.r : (ra == 0 ? -uab.i : uab.i);
      }
      if (k < 6 && (uint32_t)i < LS) bv = B[k * 16 + i];
      d = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, d, 0, 0, 0);
    }
    float *D = dl + (site * 2 + sp) * kMfmaRows * 16;
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int m = 4 * q + v;
      if (m < 6 && (uint32_t)i < LS) D[m * 16 + i] = d[v];
    }
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float *d0 = dl + ((hl * 2 + 0) * kMfmaRows + 2 * a) * 16 + s;
    const float *d1 = dl + ((hl * 2 + 1) * kMfmaRows + 2 * a) * 16 + s;
    recon_colour<float, L::MU, L::PLUS>(acc, a, cx<float>{d0[0], d0[16]}, cx<float>{d1[0], d1[16]});
  }
}

template <bool DAG, int EPI>
__global__ void __launch_bounds__(64 * kGenMaxWaves)
    k_dhop_mfma(const cx<float> *__restrict__ in, cx<float> *__restrict__ out, const cx<float> *__restrict__ U,
                KGeom g, int cb_out, const Gate *__restrict__ gate, uint32_t ntiles, int map, int flags,
                HopEpi<float> ep, GShape sh) {
  if (gated(gate)) return;
  constexpr bool PHI = (EPI & 1) != 0;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const uint32_t spb = sh.SPB;
  cx<float> *links = reinterpret_cast<cx<float> *>(lds_raw);
  // the products of a (site, spin) overwrite its staged half-spinors: the wave that reads them writes
  // them, and a lane reads back exactly the entries it staged
  float *hsl = reinterpret_cast<float *>(links + spb * kLinkPad);
  float *dl = hsl;
  const bool sweep = map == 4;
  const uint32_t tile = tile_of_block(blockIdx.x, gridDim.x, sweep ? 1 : map);
  if (tile >= ntiles) return;  // uniform across the workgroup
  const uint32_t H4 = (uint32_t)g.H4, LS = sh.LS;
  const uint32_t t = threadIdx.x;
  const uint32_t hl = fdiv(t, sh.fLS), s = t - hl * LS;
  const uint32_t p0 = tile * spb;
  const bool valid = p0 + hl < H4;
  const uint32_t pos = valid ? p0 + hl : p0;
  const uint32_t h = sweep ? sweep_site(g, pos) : pos;
  uint32_t nb[8];
  site_neighbours(g, h, cb_out, nb);
  for (uint32_t pk = s; pk < 72; pk += LS) links[hl * kLinkPad + pk] = ldc<float>(U, pk * H4 + h);
  cx<float> f[12], acc[12];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0.f, 0.f};
  // every lane takes part in the staging and the products (lanes past the lattice read their
  // tile's first site and store nothing); the next leg's neighbour load is issued before the
  // products of this one
  load_site<float>(f, in, leg_idx<DAG, 0, 0>(nb, LS, s));
  __syncthreads();
#define GX_MFMA_LEG(L)                                                                           \
  {                                                                                              \
    cx<float> fc[12];                                                                            \
    _Pragma("unroll") for (int k = 0; k < 12; k++) fc[k] = f[k];                                 \
    if (L < 7) load_site<float>(f, in, leg_idx<DAG, 0, (L < 7 ? L + 1 : 7)>(nb, LS, s));         \
    mfma_leg<DAG, L>(acc, fc, hsl, dl, links, hl, s, LS, spb);                                   \
  }
  GX_MFMA_LEG(0) GX_MFMA_LEG(1) GX_MFMA_LEG(2) GX_MFMA_LEG(3) GX_MFMA_LEG(4) GX_MFMA_LEG(5) GX_MFMA_LEG(6)
  GX_MFMA_LEG(7)
#undef GX_MFMA_LEG
  if (!valid) return;
  const uint32_t o = h * LS + s;
  if (EPI != 0) {
    const float al = ep.coef[s], be = ep.coef[LS + s];  // EPI 1
#pragma unroll
    for (int k = 0; k < 12; k++) acc[k] = {al * acc[k].r, al * acc[k].i};
    if (PHI) axpy_site<float>(acc, be, ep.phi, o);
  }
  store_site<float>(out, acc, o, (flags & 1) != 0);
}

// k_smat in the general tile shape: out = A(in) (+ beta phi) and optionally out2 = C(in), with A and
// C block diagonal in groups of LB slices.
template <typename S, bool OUT2, bool PHI>
__global__ void __launch_bounds__(64 * kGenMaxWaves)
    k_smat_g(const cx<S> *__restrict__ in, const cx<S> *__restrict__ phi, typename Sto<S>::C beta,
             cx<S> *__restrict__ out, const typename Sto<S>::C *__restrict__ mats, cx<S> *__restrict__ out2,
             const typename Sto<S>::C *__restrict__ mats2, uint32_t H4, const Gate *__restrict__ gate, GShape sh) {
  typedef typename Sto<S>::C T;
  if (gated(gate)) return;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  cx<T> *lds = reinterpret_cast<cx<T> *>(lds_raw);
  const uint32_t LS = sh.LS, epb = blockDim.x, t = threadIdx.x;
  const uint32_t hl = fdiv(t, sh.fLS), s = t - hl * LS;
  const uint32_t h = blockIdx.x * sh.SPB + hl;
  const bool valid = h < H4;
  const uint32_t o = h * LS + s;
  cx<T> f[12], f2[12];
  if (valid) {
    load_site<S>(f, in, o);
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) f[k] = {0, 0};
  }
  if (OUT2)
#pragma unroll
    for (int k = 0; k < 12; k++) f2[k] = f[k];
  group_matrix<T>(f, lds, mats, sh, t, hl, s, epb);
  if (OUT2) group_matrix<T>(f2, lds, mats2, sh, t, hl, s, epb);
  if (!valid) return;
  if (PHI) axpy_site<S>(f, beta, phi, o);
  store_site<S>(out, f, o, false);
  if (OUT2) store_site<S>(out2, f2, o, false);
}

// Waves per workgroup of the general kernels for extent LS (0: the templated kernels, -1: no
// shape fits): a multiple of LS / gcd(LS, 64), at most kGenMaxWaves, with the link tile (and the
// matrix staging) within 64 KB of LDS.  knob > 0 asks for that many waves when it is a valid
// shape; knob 0 keeps the templated kernels for LS <= 16 and otherwise takes the smallest valid
// shape of at least 4 waves (or the largest below).
static int general_waves(int LS, int knob, size_t csz, bool mat, int lb) {
  int gcd = 64, a = LS;
  while (a) {
    const int r = gcd % a;
    gcd = a;
    a = r;
  }
  const int w0 = LS / gcd;
  auto fits = [&](int w) {
    const size_t spb = (size_t)64 * w / LS;
    const size_t links = spb * kLinkPad * csz, stage = mat ? (size_t)6 * 64 * w * csz : 0;
    return w % w0 == 0 && w <= kGenMaxWaves && std::max(links, stage) <= 65536;
  };
  const bool templated_ok = LS <= kMaxLs && lb == LS;
  // the general hop stages at most kGenLinkPerThread link entries per thread: LS >= kGenMinLs
  if (knob > 0 && LS >= kGenMinLs && fits(knob)) return knob;
  if (knob < 0) {  // the smallest whole-site shape of at least 4 waves, for any LS
    for (int w = w0; w <= kGenMaxWaves; w += w0)
      if (w >= 4 && fits(w)) return w;
    for (int w = kGenMaxWaves / w0 * w0; w >= w0; w -= w0)
      if (fits(w)) return w;
    return -1;
  }
  if (LS < kGenMinLs) return templated_ok ? 0 : -1;
  if (knob == 0 && templated_ok) return 0;
  int best = -1;
  for (int w = w0; w <= kGenMaxWaves; w += w0)
    if (fits(w)) {
      best = w;
      if (w >= 4) break;
    }
  return best < 0 && templated_ok ? 0 : best;
}

// Workgroups of `threads` threads and `lds` bytes of LDS that the device holds at once, rounded down
// to a multiple of 8 (the XCD count): 2 waves per SIMD, the occupancy of the hop kernels, bounded by
// LDS (160 KB per CU).
static int resident_groups(int threads, size_t lds) {
  static int cus = 0;
  if (!cus) {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    cus = prop.multiProcessorCount;
  }
  const int waves = threads / 64;
  int per_cu = std::max(1, 8 / waves);  // 4 SIMDs x 2 waves
  if (lds > 0) per_cu = std::min<int>(per_cu, (int)(160 * 1024 / lds));
  return std::max(1, per_cu) * cus / 8 * 8;
}

static GShape make_shape(int LS, int LB, int W) {
  GShape sh;
  sh.LS = (uint32_t)LS;
  sh.LB = (uint32_t)LB;
  sh.SPB = (uint32_t)(64 * W / LS);
  sh.fLS = make_fdiv((uint32_t)LS);
  sh.fLB = make_fdiv((uint32_t)LB);
  return sh;
}

#ifndef GX_LS_CASES
#define GX_LS_CASES(M) \
  M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9) M(10) M(11) M(12) M(13) M(14) M(15) M(16)
#endif

// epi: 0 out = hop; 1 out = alpha hop + beta phi; 2 out = alpha A(hop) [+ beta phi when phi is
// given] (A: mat).  fz: reductions in the epilogue of the 4-wave whole-site kernel (HopFuse).
template <typename S>
int launch_dhop_epi(const Geom &g, const void *in, void *out, const void *U, int cb_out, int dag, const Gate *gate,
                    hipStream_t st, int epi, const void *phi, const void *mat, double alpha, double beta, int lb,
                    const void *coef, const HopFuse *fz) {
  typedef typename Sto<S>::C T;
  if (fz && fz->nparts) *fz->nparts = 0;
  if (lb <= 0) lb = g.Ls;
  if (13 * 2 * g.half_pad() >= (int64_t)1 << 31) {
    set_error("dhop: lattice too large for 32-bit site indexing");
    return GX_EINVAL;
  }
  if (epi == 2 && g.Ls < 2) {
    set_error("dhop: matrix epilogue needs Ls >= 2");
    return GX_EINVAL;
  }
  if ((epi == 1 && (!phi || !coef)) || (epi == 2 && !mat)) {
    set_error("dhop: epilogue operand missing");
    return GX_EINVAL;
  }
  const int kepi = (epi == 2 && phi) ? 3 : epi;
  int map = tune_get(GX_TUNE_DHOP_MAP, prec_of<S>());
  KGeom k = kgeom(g);
  if (map == 4) {
    // blocks of Lxh x by x bz sites per t-slice with three t-slices of the input (and the links)
    // within 3.5 MB of an XCD's L2; halve by or bz while they are even and the block is too big
    const double site_bytes = (double)g.Ls * 12 * sizeof(cx<S>) + 72.0 * sizeof(cx<S>);
    int by = g.L[1], bz = g.L[2];
    while (3.0 * k.Lxh * by * bz * site_bytes > 3.5e6) {
      if (bz >= by && bz % 2 == 0) bz /= 2;
      else if (by % 2 == 0) by /= 2;
      else if (bz % 2 == 0) bz /= 2;
      else break;
    }
    set_sweep_blocks(k, by, bz);
  }
  const bool pf = tune_get(GX_TUNE_DHOP_PF, prec_of<S>()) != 0;
  const int flags = tune_get(GX_TUNE_DHOP_NT, prec_of<S>()) ? 1 : 0;
  const bool ord = epi == 0 && tune_get(GX_TUNE_DHOP_ORDER, prec_of<S>()) != 0 && (g.Ls == 1 || g.Ls == 12 || g.Ls == 16);
  const cx<S> *pin = (const cx<S> *)in;
  cx<S> *pout = (cx<S> *)out;
  const cx<S> *pU = (const cx<S> *)U;
  HopEpi<S> ep{(const cx<S> *)phi, (const T *)mat, (T)alpha, (T)beta, (const T *)coef,
               fz ? fz->npart : nullptr,  (cx<S> *)(fz ? fz->r : nullptr),  fz ? fz->a : nullptr,
               fz ? fz->astride : 1};
  if constexpr (std::is_same<S, float>::value) {
    // the link products on the matrix cores: FP32 multi-RHS (Wilson, no matrix epilogue), Ls <= 16
    if (tune_get(GX_TUNE_DHOP_MFMA, prec_of<S>()) && kepi <= 1 && g.Ls >= 2 && g.Ls <= 16 && lb == g.Ls) {
      const int Wm = general_waves(g.Ls, -1, sizeof(cx<T>), false, lb);  // smallest shape of >= 4 waves
      if (Wm > 0) {
        const GShape sh = make_shape(g.Ls, lb, Wm);
        const uint32_t ntiles = (uint32_t)((g.H4 + sh.SPB - 1) / sh.SPB);
        const uint32_t pad = (map == 1 || map == 4) ? 8 : (map == 2 ? 64 : (map == 3 ? 256 : 1));
        const uint32_t blocks = (ntiles + pad - 1) / pad * pad;
        const size_t lds = (size_t)sh.SPB * kLinkPad * sizeof(cx<float>) + (size_t)sh.SPB * 2 * kMfmaRows * 16 * 4;
        if (lds <= 65536) {
#define GX_MFMA_GO(DAG, EPI)                                                                                     \
  hipLaunchKernelGGL((k_dhop_mfma<DAG, EPI>), dim3(blocks), dim3(64 * Wm), lds, st, pin, pout, pU, k, cb_out, gate, \
                     ntiles, map, flags, ep, sh)
          if (dag) {
            if (kepi == 1) GX_MFMA_GO(true, 1); else GX_MFMA_GO(true, 0);
          } else {
            if (kepi == 1) GX_MFMA_GO(false, 1); else GX_MFMA_GO(false, 0);
          }
#undef GX_MFMA_GO
          GX_CHECK_HIP(hipGetLastError());
          return GX_OK;
        }
      }
    }
  }
  // row tiles: 2 x 2 rows of x in LDS (whole tiles only: Ly and Lz are even)
  if (tune_get(GX_TUNE_DHOP_TILE, prec_of<S>()) && g.Ls % lb == 0) {
    const uint32_t RW = (uint32_t)(k.Lxh * g.Ls), nthr = 4 * RW;
    const size_t lds = ((size_t)12 * nthr + (size_t)4 * k.Lxh * kLinkPad) * sizeof(cx<T>) +
                       (Sto<S>::H ? (size_t)nthr * sizeof(T) : 0);  // FP16: the entries' site scales
    if (nthr <= (uint32_t)rows_max_threads<S>() && lds <= 160 * 1024 && g.L[1] >= 4 && g.L[2] >= 4) {
      RowTile rt;
      rt.LS = (uint32_t)g.Ls;
      rt.RW = RW;
      rt.nyb = (uint32_t)g.L[1] / 2;
      rt.fLS = make_fdiv(rt.LS);
      rt.fRW = make_fdiv(RW);
      rt.fLt = make_fdiv((uint32_t)g.L[3]);
      rt.fnyb = make_fdiv(rt.nyb);
      rt.fQ = make_fdiv((uint32_t)(4 * k.Lxh));
      const GShape sh = make_shape(g.Ls, lb, 1);  // LS and the matrix groups for the epilogue
      const uint32_t nzb = (uint32_t)g.L[2] / 2, ntiles = (uint32_t)g.L[3] * rt.nyb * nzb;
      rt.xblk = 0;
      if (map == 4) {
        // XCD boxes: the split pY x pZ x pT = 8 (each dividing its extent) with the fewest face rows
        // per box, counting a t face as a whole box face when two t-slices of the box's input and
        // links exceed 3.5 MB of the XCD's L2 (then the t neighbours are not reused)
        const double row_bytes = (double)k.Lxh * (g.Ls * 12.0 + 72.0) * sizeof(cx<S>);
        double best = 1e30;
        for (uint32_t pY = 1; pY <= 8; pY *= 2)
          for (uint32_t pZ = 1; pY * pZ <= 8; pZ *= 2) {
            const uint32_t pT = 8 / (pY * pZ);
            if (rt.nyb % pY || nzb % pZ || (uint32_t)g.L[3] % pT) continue;
            const double ry = 2.0 * rt.nyb / pY, rz = 2.0 * nzb / pZ, tl = (double)g.L[3] / pT;
            const bool treuse = 2.0 * ry * rz * row_bytes <= 3.5e6;
            const double face = (pY > 1 ? 2.0 / ry : 0.0) + (pZ > 1 ? 2.0 / rz : 0.0) +
                                (treuse ? (pT > 1 ? 2.0 / tl : 0.0) : 2.0);
            if (face < best) {
              best = face;
              rt.pY = pY;
              rt.pZ = pZ;
              rt.bY = rt.nyb / pY;
              rt.bZ = nzb / pZ;
              rt.tper = (uint32_t)g.L[3] / pT;
              rt.xblk = 1;
            }
          }
        if (rt.xblk) {
          rt.per = ntiles / 8;
          rt.nc = rt.bY * rt.bZ;
          rt.fper = make_fdiv(rt.per);
          rt.fnc = make_fdiv(rt.nc);
          rt.fbY = make_fdiv(rt.bY);
          rt.fpY = make_fdiv(rt.pY);
          rt.fpZ = make_fdiv(rt.pZ);
        }
      }
      const int tmap = map == 4 ? 1 : map;  // map 4: the XCD boxes above, one map-1 slab each
      const uint32_t pad = tmap == 1 ? 8 : (tmap == 2 ? 64 : (tmap == 3 ? 256 : 1));
      const uint32_t blocks = (ntiles + pad - 1) / pad * pad;
#define GX_ROWS_GO(DAG, EPI)                                                                                    \
  do {                                                                                                           \
    GX_CHECK_HIP(hipFuncSetAttribute((const void *)k_dhop_rows<S, DAG, EPI>,                                   \
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                    \
    hipLaunchKernelGGL((k_dhop_rows<S, DAG, EPI>), dim3(blocks), dim3(nthr), lds, st, pin, pout, pU, k, cb_out, \
                       gate, ntiles, tmap, flags, ep, sh, rt);                                                   \
  } while (0)
#define GX_ROWS_EPI(EPI) \
  if (dag) GX_ROWS_GO(true, EPI); else GX_ROWS_GO(false, EPI)
      switch (kepi) {
        case 0: GX_ROWS_EPI(0); break;
        case 1: GX_ROWS_EPI(1); break;
        case 2: GX_ROWS_EPI(2); break;
        default: GX_ROWS_EPI(3); break;
      }
#undef GX_ROWS_EPI
#undef GX_ROWS_GO
      GX_CHECK_HIP(hipGetLastError());
      return GX_OK;
    }
  }
  const int W = general_waves(g.Ls, tune_get(GX_TUNE_DHOP_SHAPE, prec_of<S>()), sizeof(cx<T>), kepi >= 2, lb);
  if (W < 0 || g.Ls % lb != 0) {
    set_error("dhop: no tile shape for Ls=%d (matrix group %d): Ls > 16 needs an odd part of at most 7", g.Ls, lb);
    return GX_EINVAL;
  }
  if (W > 0) {
    const GShape sh = make_shape(g.Ls, lb, W);
    const uint32_t ntiles = (uint32_t)((g.H4 + sh.SPB - 1) / sh.SPB);
    const uint32_t pad = (map == 1 || map == 4) ? 8 : (map == 2 ? 64 : (map == 3 ? 256 : 1));
    const uint32_t blocks = (ntiles + pad - 1) / pad * pad;
    const size_t lds = std::max((size_t)sh.SPB * kLinkPad, kepi >= 2 ? (size_t)6 * 64 * W : (size_t)0) * sizeof(cx<T>);
    // persistent grid: the resident capacity of the device (multiple of 8), when smaller than the tiles
    const int persist_wg = tune_get(GX_TUNE_DHOP_PERSIST, prec_of<S>()) ? resident_groups(64 * W, lds) : 0;
    const int persist = persist_wg > 0 && (uint32_t)persist_wg < blocks ? 1 : 0;
    const uint32_t grid = persist ? (uint32_t)persist_wg : blocks;
#define GX_DHOPG_GO(DAG, EPI)                                                                                        \
  hipLaunchKernelGGL((k_dhop_g<S, DAG, EPI>), dim3(grid), dim3(64 * W), lds, st, pin, pout, pU, k, cb_out, gate,     \
                     ntiles, map, flags, ep, sh, blocks, persist)
#define GX_DHOPG_EPI(EPI) \
  if (dag) GX_DHOPG_GO(true, EPI); else GX_DHOPG_GO(false, EPI)
    // fused reductions: EPI 5 (kepi 1, dag 0), 6 (kepi 1, dag 1, r), 7 (kepi 3, dag 1, r); one
    // workgroup per tile only (the persistent grid runs the separate passes)
    const bool gnorm = !persist && fz && fz->npart && fz->nparts &&
                       ((kepi == 1 && (fz->r ? (dag && fz->a) : !dag)) || (kepi == 3 && dag && fz->r && fz->a));
    if (gnorm) {
      *fz->nparts = ntiles;
      if (!dag) GX_DHOPG_GO(false, 5);
      else if (kepi == 3) GX_DHOPG_GO(true, 7);
      else GX_DHOPG_GO(true, 6);
    } else {
      switch (kepi) {
        case 0: GX_DHOPG_EPI(0); break;
        case 1: GX_DHOPG_EPI(1); break;
        case 2: GX_DHOPG_EPI(2); break;
        default: GX_DHOPG_EPI(3); break;
      }
    }
#undef GX_DHOPG_EPI
#undef GX_DHOPG_GO
    GX_CHECK_HIP(hipGetLastError());
    return GX_OK;
  }
  const int spb = 4 * (64 / g.Ls);
  const uint32_t ntiles = (uint32_t)((g.H4 + spb - 1) / spb);
  const uint32_t pad = (map == 1 || map == 4) ? 8 : (map == 2 ? 64 : (map == 3 ? 256 : 1));
  const uint32_t blocks = (ntiles + pad - 1) / pad * pad;
  // EPI 5: kepi 1, dag 0; EPI 6: kepi 1, dag 1, r; EPI 7: kepi 3 (matrix and phi), dag 1, r
  const bool norm = fz && fz->npart && fz->nparts && g.Ls >= 2 &&
                    ((kepi == 1 && (fz->r ? (dag && fz->a) : !dag)) || (kepi == 3 && dag && fz->r && fz->a));
  if (norm) *fz->nparts = ntiles;
#define GX_DHOP_GO(LS, DAG, PF, ORD, EPI)                                                                         \
  hipLaunchKernelGGL((k_dhop<S, LS, DAG, PF, ORD, EPI>), dim3(blocks), dim3(256), 0, st, pin, pout, pU, k, cb_out, \
                     gate, ntiles, map, flags, ep)
#define GX_DHOP_CASE_O(LS, ORD, EPI)                   \
  if (dag) {                                           \
    if (pf) GX_DHOP_GO(LS, true, true, ORD, EPI);      \
    else GX_DHOP_GO(LS, true, false, ORD, EPI);        \
  } else {                                             \
    if (pf) GX_DHOP_GO(LS, false, true, ORD, EPI);     \
    else GX_DHOP_GO(LS, false, false, ORD, EPI);       \
  }
#define GX_DHOP_CASE(LS)                                         \
  case LS:                                                       \
    if (norm) {                                                  \
      if constexpr (LS >= 2) {                                   \
        if (dag && kepi == 3) {                                  \
          if (pf) GX_DHOP_GO(LS, true, true, 0, 7);              \
          else GX_DHOP_GO(LS, true, false, 0, 7);                \
        } else if (dag) {                                        \
          if (pf) GX_DHOP_GO(LS, true, true, 0, 6);              \
          else GX_DHOP_GO(LS, true, false, 0, 6);                \
        } else {                                                 \
          if (pf) GX_DHOP_GO(LS, false, true, 0, 5);             \
          else GX_DHOP_GO(LS, false, false, 0, 5);               \
        }                                                        \
      }                                                          \
    } else if (kepi == 1) {                                      \
      GX_DHOP_CASE_O(LS, 0, 1)                                   \
    } else if (kepi == 2) {                                      \
      if constexpr (LS >= 2) {                                   \
        GX_DHOP_CASE_O(LS, 0, 2)                                 \
      }                                                          \
    } else if (kepi == 3) {                                      \
      if constexpr (LS >= 2) {                                   \
        GX_DHOP_CASE_O(LS, 0, 3)                                 \
      }                                                          \
    } else if (ord && (LS == 1 || LS == 12 || LS == 16)) {       \
      if constexpr (LS == 1 || LS == 12 || LS == 16) {           \
        GX_DHOP_CASE_O(LS, 1, 0)                                 \
      }                                                          \
    } else {                                                     \
      GX_DHOP_CASE_O(LS, 0, 0)                                   \
    }                                                            \
    break;
  switch (g.Ls) {
    GX_LS_CASES(GX_DHOP_CASE)
    default:
      set_error("dhop: Ls=%d unsupported (1..%d)", g.Ls, kMaxLs);
      return GX_EINVAL;
  }
#undef GX_DHOP_CASE
#undef GX_DHOP_CASE_O
#undef GX_DHOP_GO
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

template <typename S>
int launch_dhop(const Geom &g, const void *in, void *out, const void *U, int cb_out, int dag, const Gate *gate,
                hipStream_t st) {
  return launch_dhop_epi<S>(g, in, out, U, cb_out, dag, gate, st, 0, nullptr, nullptr, 1.0, 0.0, 0, nullptr, nullptr);
}

// out_s = sum_t A_chi[s][t] in_t (+ beta phi_s with PHI), and with OUT2 also
// out2_s = sum_t C_chi[s][t] in_t from the same staged input.  mats: [2][LS][LS] in T,
// chirality 0 = spins 0,1.  One lane per (h, s); the site's s-row is staged in LDS.
template <typename S, int LS, bool OUT2, bool PHI>
__global__ void __launch_bounds__(256) k_smat(const cx<S> *__restrict__ in, const cx<S> *__restrict__ phi,
                                              typename Sto<S>::C beta, cx<S> *__restrict__ out,
                                              const typename Sto<S>::C *__restrict__ mats, cx<S> *__restrict__ out2,
                                              const typename Sto<S>::C *__restrict__ mats2, int64_t H4,
                                              int64_t stride, const Gate *__restrict__ gate) {
  typedef typename Sto<S>::C T;
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
  {
    cx<T> f[12];
    if (valid) {
      load_site<S>(f, in, o);
    } else {
#pragma unroll
      for (int k = 0; k < 12; k++) f[k] = {0, 0};
    }
#pragma unroll
    for (int k = 0; k < 12; k++) lds[w][k][lane] = f[k];
  }
  __syncthreads();
  if (!valid) return;
  cx<T> acc[12], acc2[OUT2 ? 12 : 1];
#pragma unroll
  for (int k = 0; k < 12; k++) acc[k] = {0, 0};
  if (OUT2)
#pragma unroll
    for (int k = 0; k < 12; k++) acc2[k] = {0, 0};
#pragma unroll
  for (int t = 0; t < LS; t++) {
    const T ap = mats[s * LS + t], am = mats[LS * LS + s * LS + t];
    T cp = 0, cm = 0;
    if (OUT2) {
      cp = mats2[s * LS + t];
      cm = mats2[LS * LS + s * LS + t];
    }
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const cx<T> v = lds[w][k][ls * LS + t];
      const T a = k < 6 ? ap : am;
      acc[k].r = fma(a, v.r, acc[k].r);
      acc[k].i = fma(a, v.i, acc[k].i);
      if (OUT2) {
        const T c = k < 6 ? cp : cm;
        acc2[k].r = fma(c, v.r, acc2[k].r);
        acc2[k].i = fma(c, v.i, acc2[k].i);
      }
    }
  }
  if (PHI) axpy_site<S>(acc, beta, phi, o);
  store_site<S>(out, acc, o, false);
  if (OUT2) store_site<S>(out2, acc2, o, false);
}

template <typename S>
int launch_smat2(const Geom &g, const void *in, const void *phi, double beta, void *out, const void *mats, void *out2,
                 const void *mats2, const Gate *gate, hipStream_t st, int lb) {
  typedef typename Sto<S>::C T;
  if (lb <= 0) lb = g.Ls;
  if (g.Ls > kMaxLs || lb != g.Ls) {
    const int W = general_waves(g.Ls, 0, sizeof(cx<T>), true, lb);
    if (W <= 0 || g.Ls % lb != 0) {
      set_error("smat: no tile shape for Ls=%d (matrix group %d)", g.Ls, lb);
      return GX_EINVAL;
    }
    const GShape sh = make_shape(g.Ls, lb, W);
    const uint32_t blocks = (uint32_t)((g.H4 + sh.SPB - 1) / sh.SPB);
    const size_t lds = (size_t)6 * 64 * W * sizeof(cx<T>);
    const cx<S> *pin = (const cx<S> *)in, *pphi = (const cx<S> *)phi;
    cx<S> *pout = (cx<S> *)out, *pout2 = (cx<S> *)out2;
    const T *pm = (const T *)mats, *pm2 = (const T *)mats2;
#define GX_SMATG_GO(OUT2, PHI)                                                                                       \
  hipLaunchKernelGGL((k_smat_g<S, OUT2, PHI>), dim3(blocks), dim3(64 * W), lds, st, pin, pphi, (T)beta, pout, pm,   \
                     pout2, pm2, (uint32_t)g.H4, gate, sh)
    if (out2) {
      if (phi) GX_SMATG_GO(true, true);
      else GX_SMATG_GO(true, false);
    } else {
      if (phi) GX_SMATG_GO(false, true);
      else GX_SMATG_GO(false, false);
    }
#undef GX_SMATG_GO
    GX_CHECK_HIP(hipGetLastError());
    return GX_OK;
  }
  const int spw = 64 / g.Ls;
  const int64_t waves = (g.H4 + spw - 1) / spw;
  const int64_t blocks = (waves * 64 + 255) / 256;
  const int64_t stride = g.half_pad();
  const cx<S> *pin = (const cx<S> *)in;
  const cx<S> *pphi = (const cx<S> *)phi;
  cx<S> *pout = (cx<S> *)out, *pout2 = (cx<S> *)out2;
  const T *pm = (const T *)mats, *pm2 = (const T *)mats2;
#define GX_SMAT_GO(LS, OUT2, PHI)                                                                           \
  hipLaunchKernelGGL((k_smat<S, LS, OUT2, PHI>), dim3(blocks), dim3(256), 0, st, pin, pphi, (T)beta, pout, pm, \
                     pout2, pm2, g.H4, stride, gate)
#define GX_SMAT_CASE(LS)                    \
  case LS:                                  \
    if (out2) {                             \
      if (phi) GX_SMAT_GO(LS, true, true);  \
      else GX_SMAT_GO(LS, true, false);     \
    } else {                                \
      if (phi) GX_SMAT_GO(LS, false, true); \
      else GX_SMAT_GO(LS, false, false);    \
    }                                       \
    break;
  switch (g.Ls) {
    GX_LS_CASES(GX_SMAT_CASE)
    default:
      set_error("smat: Ls=%d unsupported", g.Ls);
      return GX_EINVAL;
  }
#undef GX_SMAT_CASE
#undef GX_SMAT_GO
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

template <typename S>
int launch_smat(const Geom &g, const void *in, const void *phi, void *out, const void *mats, const Gate *gate,
                hipStream_t st, int lb) {
  return launch_smat2<S>(g, in, phi, 1.0, out, mats, nullptr, nullptr, gate, st, lb);
}

// ---------------- BLAS1 ----------------
static int64_t stream_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return b < 1 ? 1 : (b > 65536 ? 65536 : b);
}

// z = a x + b y (y optional) over nsites sites of half or full fields (stride sites per half).
// FP64/FP32 stream the entries; FP16 works per site (its stores need the site's scale).
template <typename S>
__global__ void __launch_bounds__(256) k_axpby(int64_t nsites, int64_t stride, typename Sto<S>::C a,
                                               const cx<S> *__restrict__ x, typename Sto<S>::C b,
                                               const cx<S> *__restrict__ y, cx<S> *__restrict__ z,
                                               const Gate *__restrict__ gate) {
  typedef typename Sto<S>::C T;
  if (gated(gate)) return;
  if constexpr (!Sto<S>::H) {
    const int64_t n = 12 * nsites;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
      const cx<T> xv = x[i];
      if (y) {
        const cx<T> yv = y[i];
        z[i] = {fma(a, xv.r, b * yv.r), fma(a, xv.i, b * yv.i)};
      } else {
        z[i] = {a * xv.r, a * xv.i};
      }
    }
  } else {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites;
         j += (int64_t)gridDim.x * blockDim.x) {
      const int64_t hf = j / stride, i = j - hf * stride, base = hf * Sto<S>::NCOL * stride;
      cx<T> xv[12], yv[12];
      load_site<S>(xv, x + base, i);
      if (y) {
        load_site<S>(yv, y + base, i);
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {fma(a, xv[k].r, b * yv[k].r), fma(a, xv[k].i, b * yv[k].i)};
      } else {
#pragma unroll
        for (int k = 0; k < 12; k++) xv[k] = {a * xv[k].r, a * xv[k].i};
      }
      store_site<S>(z + base, xv, i, false);
    }
  }
}

template <typename S>
int launch_axpby(int64_t nsites, int64_t stride, double a, const void *x, double b, const void *y, void *z,
                 const Gate *gate, hipStream_t st) {
  typedef typename Sto<S>::C T;
  const int64_t work = Sto<S>::H ? nsites : 12 * nsites;
  hipLaunchKernelGGL((k_axpby<S>), dim3(stream_blocks(work)), dim3(256), 0, st, nsites, stride, (T)a,
                     (const cx<S> *)x, (T)b, (const cx<S> *)y, (cx<S> *)z, gate);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// out_e = c[e % Ls] in_e over the (h, s) entries of a half field: a diagonal operator with one
// coefficient per fifth-dimension slice (padding entries are left as they are: zero)
template <typename S>
__global__ void __launch_bounds__(256) k_scale_slices(int64_t nent, int Ls, const cx<S> *__restrict__ in,
                                                      cx<S> *__restrict__ out, const typename Sto<S>::C *__restrict__ c,
                                                      const Gate *__restrict__ gate) {
  typedef typename Sto<S>::C T;
  if (gated(gate)) return;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += (int64_t)gridDim.x * blockDim.x) {
    const T a = c[e % Ls];
    cx<T> v[12];
    load_site<S>(v, in, e);
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = {a * v[k].r, a * v[k].i};
    store_site<S>(out, v, e, false);
  }
}

template <typename S>
int launch_scale_slices(const Geom &g, const void *in, void *out, const void *c, const Gate *gate, hipStream_t st) {
  typedef typename Sto<S>::C T;
  const int64_t nent = g.H4 * g.Ls;
  hipLaunchKernelGGL((k_scale_slices<S>), dim3(stream_blocks(nent)), dim3(256), 0, st, nent, g.Ls, (const cx<S> *)in,
                     (cx<S> *)out, (const T *)c, gate);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// flat element conversion (links, and FP64 <-> FP32 fields)
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

// per-site field conversion between storages (any pair that involves FP16)
template <typename Sin, typename Sout>
__global__ void __launch_bounds__(256) k_convert_site(int64_t nsites, int64_t stride, const cx<Sin> *__restrict__ in,
                                                      cx<Sout> *__restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride;
    cx<typename Sto<Sin>::C> a[12];
    cx<typename Sto<Sout>::C> b[12];
    load_site<Sin>(a, in + hf * Sto<Sin>::NCOL * stride, i);
#pragma unroll
    for (int k = 0; k < 12; k++) b[k] = {(typename Sto<Sout>::C)a[k].r, (typename Sto<Sout>::C)a[k].i};
    store_site<Sout>(out + hf * Sto<Sout>::NCOL * stride, b, i, false);
  }
}

// out += in over the sites of an FP64 field (defect correction folding an inner solution into the
// outer one: the sum of a conversion then an axpby with unit coefficients, in one pass)
template <typename Sin>
__global__ void __launch_bounds__(256) k_accum_site(int64_t nsites, int64_t stride, const cx<Sin> *__restrict__ in,
                                                    cx<double> *__restrict__ out) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride;
    cx<typename Sto<Sin>::C> a[12];
    cx<double> b[12];
    load_site<Sin>(a, in + hf * Sto<Sin>::NCOL * stride, i);
    load_site<double>(b, out + hf * 12 * stride, i);
#pragma unroll
    for (int k = 0; k < 12; k++) b[k] = {b[k].r + (double)a[k].r, b[k].i + (double)a[k].i};
    store_site<double>(out + hf * 12 * stride, b, i, false);
  }
}

template <typename Sin>
int launch_accum_site(int64_t nsites, int64_t stride, const void *in, void *out, hipStream_t st) {
  hipLaunchKernelGGL((k_accum_site<Sin>), dim3(stream_blocks(nsites)), dim3(256), 0, st, nsites, stride,
                     (const cx<Sin> *)in, (cx<double> *)out);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

template <typename Sin, typename Sout>
int launch_convert_site(int64_t nsites, int64_t stride, const void *in, void *out, hipStream_t st) {
  hipLaunchKernelGGL((k_convert_site<Sin, Sout>), dim3(stream_blocks(nsites)), dim3(256), 0, st, nsites, stride,
                     (const cx<Sin> *)in, (cx<Sout> *)out);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- deterministic reductions ----------------
// Two kernels: every block of the streaming kernel writes its partial sum (wave shuffles, then LDS
// across the 4 waves), and a one-block kernel adds the partials in block order.  The kernel
// boundary makes the partials visible to the second kernel, so no device-scope fence or atomic
// is needed (on a multi-XCD part such a fence writes back the XCD's L2).  The grid size depends
// only on the problem size, so results are bitwise reproducible.
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

template <int NV>
__device__ __forceinline__ void write_partial(const double (&v)[NV], double *partial) {
  if (threadIdx.x == 0)
#pragma unroll
    for (int j = 0; j < NV; j++) partial[blockIdx.x * NV + j] = v[j];
}

// sum of nparts partials (NV values each) in block order, by one block of 256 threads
template <int NV>
__device__ __forceinline__ void sum_partials(double (&v)[NV], const double *partial, int nparts, double (*sh)[4]) {
#pragma unroll
  for (int j = 0; j < NV; j++) v[j] = 0;
  for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
    for (int j = 0; j < NV; j++) v[j] += partial[b * NV + j];
  block_sum<NV>(v, sh);
}

// entry j of a (possibly full) field: half hf = j / stride (stride = padded entries per half),
// entry i = j % stride of that half, whose block starts hf * NCOL * stride entries in.
template <typename S, int MODE>
__global__ void __launch_bounds__(256) k_reduce(int64_t nsites, int64_t stride, const cx<S> *__restrict__ x,
                                                const cx<S> *__restrict__ y, cx<S> *__restrict__ z,
                                                typename Sto<S>::C a, double *partial) {
  typedef typename Sto<S>::C T;
  __shared__ double sh[2][4];
  double v[2] = {0, 0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride, base = hf * Sto<S>::NCOL * stride;
    cx<T> xv[12], yv[12];
    load_site<S>(xv, x + base, i);
    if (MODE != 1) load_site<S>(yv, y + base, i);
    T sr = 0, si = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      if (MODE == 0) {
        sr = fma(xv[k].r, yv[k].r, sr);
        sr = fma(xv[k].i, yv[k].i, sr);
        si = fma(xv[k].r, yv[k].i, si);
        si = fma(-xv[k].i, yv[k].r, si);
      } else if (MODE == 1) {
        sr = fma(xv[k].r, xv[k].r, sr);
        sr = fma(xv[k].i, xv[k].i, sr);
      } else if (MODE == 3) {  // Re<x|y> and |x|^2 in one pass
        sr = fma(xv[k].r, yv[k].r, sr);
        sr = fma(xv[k].i, yv[k].i, sr);
        si = fma(xv[k].r, xv[k].r, si);
        si = fma(xv[k].i, xv[k].i, si);
      } else {
        yv[k] = {fma(a, xv[k].r, yv[k].r), fma(a, xv[k].i, yv[k].i)};
        sr = fma(yv[k].r, yv[k].r, sr);
        sr = fma(yv[k].i, yv[k].i, sr);
      }
    }
    if (MODE == 2) store_site<S>(z + base, yv, i, false);
    v[0] += (double)sr;
    v[1] += (double)si;
  }
  block_sum<2>(v, sh);
  write_partial<2>(v, partial);
}

static __global__ void __launch_bounds__(256) k_reduce_fin(const double *partial, int nparts, double *result, int nout) {
  __shared__ double sh[2][4];
  double v[2];
  sum_partials<2>(v, partial, nparts, sh);
  if (threadIdx.x == 0) {
    result[0] = v[0];
    if (nout > 1) result[1] = v[1];
  }
}

static int red_blocks(int64_t nsites) {
  int64_t b = (nsites + 255) / 256;
  if (b > kRedBlocks) b = kRedBlocks;
  return (int)(b < 1 ? 1 : b);
}

template <typename S>
int launch_reduce(Ctx *ctx, int mode, int64_t nsites, int64_t stride, const void *x, const void *y, void *z, double a,
                  int slot) {
  typedef typename Sto<S>::C T;
  const int nb = red_blocks(nsites);
  const cx<S> *px = (const cx<S> *)x;
  const cx<S> *py = (const cx<S> *)y;
  cx<S> *pz = (cx<S> *)z;
  if (mode == 0)
    hipLaunchKernelGGL((k_reduce<S, 0>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial);
  else if (mode == 1)
    hipLaunchKernelGGL((k_reduce<S, 1>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial);
  else if (mode == 3)
    hipLaunchKernelGGL((k_reduce<S, 3>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial);
  else
    hipLaunchKernelGGL((k_reduce<S, 2>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, px, py, pz, (T)a,
                       ctx->red_partial);
  hipLaunchKernelGGL(k_reduce_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb,
                     ctx->red_result + slot, (mode == 0 || mode == 3) ? 2 : 1);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- CG iteration on device scalars (ConjugateGradient.h:151-198) ----------------
// Kernels of iteration k run while k <= min(k_conv, kmax).  The norm finaliser advances k, so the
// solution update of the same iteration (launched after it) tests k-1.
__device__ __forceinline__ bool gated_prev(const CGState *st) { return st->k - 1 > st->k_conv || st->k - 1 > st->kmax; }

// partials of d = Re<p|mmp>
template <typename S>
__global__ void __launch_bounds__(256) k_cg_dot(int64_t nsites, int64_t stride, const cx<S> *__restrict__ p,
                                                const cx<S> *__restrict__ mmp, const CGState *st, double *partial) {
  typedef typename Sto<S>::C T;
  if (gated(st)) return;
  __shared__ double sh[1][4];
  double v[1] = {0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride, base = hf * Sto<S>::NCOL * stride;
    cx<T> xv[12], yv[12];
    load_site<S>(xv, p + base, i);
    load_site<S>(yv, mmp + base, i);
    T sr = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      sr = fma(xv[k].r, yv[k].r, sr);
      sr = fma(xv[k].i, yv[k].i, sr);
    }
    v[0] += (double)sr;
  }
  block_sum<1>(v, sh);
  write_partial<1>(v, partial);
}

// partials of cp = |r - a mmp|^2 with r updated in place
template <typename S>
__global__ void __launch_bounds__(256) k_cg_axpy_norm(int64_t nsites, int64_t stride, cx<S> *__restrict__ r,
                                                      const cx<S> *__restrict__ mmp, const CGState *st,
                                                      double *partial) {
  typedef typename Sto<S>::C T;
  if (gated(st)) return;
  __shared__ double sh[1][4];
  const T a = (T)(-st->a);
  double v[1] = {0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride, base = hf * Sto<S>::NCOL * stride;
    cx<T> xv[12], yv[12];
    load_site<S>(xv, mmp + base, i);
    load_site<S>(yv, r + base, i);
    T sr = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      yv[k] = {fma(a, xv[k].r, yv[k].r), fma(a, xv[k].i, yv[k].i)};
      sr = fma(yv[k].r, yv[k].r, sr);
      sr = fma(yv[k].i, yv[k].i, sr);
    }
    store_site<S>(r + base, yv, i, false);
    v[0] += (double)sr;
  }
  block_sum<1>(v, sh);
  write_partial<1>(v, partial);
}

// which = 0: d = sum, a = c / d.   which = 1: cp = sum, b = cp / c, c = cp, convergence test.
// The which = 1 finaliser advances k in every iteration, also once the solve has stopped, so the
// gate of the following solution update (k-1) closes after the last real iteration.
static __global__ void __launch_bounds__(256) k_cg_fin(const double *partial, int nparts, CGState *st, int which) {
  const bool skip = gated(st);
  if (skip && which == 0) return;
  __shared__ double sh[1][4];
  double v[1] = {0};
  if (!skip) sum_partials<1>(v, partial, nparts, sh);
  if (threadIdx.x == 0) {
    if (which == 0) {
      st->d = v[0];
      st->a = st->c / st->d;
    } else {
      if (!skip) {
        const double cp = v[0];
        st->cprev = st->c;
        st->b = cp / st->c;
        st->cp = cp;
        st->c = cp;
        if (st->delta > 0 && cp > st->maxres) st->maxres = cp;
        if (cp <= st->rsq && st->k_conv == INT_MAX) {
          st->k_conv = st->k;
        } else if (st->delta > 0 && cp < st->delta * st->maxres && st->k_conv == INT_MAX) {
          st->ru = 1;
          st->k_conv = st->k;
        } else if (st->fbtol > 0 && cp < st->fbtol && st->k_conv == INT_MAX) {
          st->ru = 2;
          st->k_conv = st->k;
        }
      }
      st->k = st->k + 1;
    }
  }
}

// psi = a p + psi; p = b p + r  (FP64/FP32 per entry, FP16 per site)
template <typename S>
__global__ void __launch_bounds__(256) k_cg_update(int64_t nsites, int64_t stride, cx<S> *__restrict__ psi,
                                                   cx<S> *__restrict__ p, const cx<S> *__restrict__ r,
                                                   const CGState *st) {
  typedef typename Sto<S>::C T;
  if (gated_prev(st)) return;
  const T a = (T)st->a, b = (T)st->b;
  const bool keep_p = st->ru == 1;  // reliable update pending: the host forms the new p
  if constexpr (!Sto<S>::H) {
    // one 16-byte group (VW complex numbers) per thread and step
    constexpr int VW = Sto<S>::VW;
    struct alignas(16) G {
      cx<T> v[VW];
    };
    G *gx_ = reinterpret_cast<G *>(psi), *gp = reinterpret_cast<G *>(p);
    const G *gr = reinterpret_cast<const G *>(r);
    const int64_t ng = 12 * nsites / VW;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < ng; q += (int64_t)gridDim.x * blockDim.x) {
      const G pv = gp[q];
      G xv = gx_[q];
#pragma unroll
      for (int j = 0; j < VW; j++) xv.v[j] = {fma(a, pv.v[j].r, xv.v[j].r), fma(a, pv.v[j].i, xv.v[j].i)};
      gx_[q] = xv;
      if (!keep_p) {
        const G rv = gr[q];
        G np;
#pragma unroll
        for (int j = 0; j < VW; j++) np.v[j] = {fma(b, pv.v[j].r, rv.v[j].r), fma(b, pv.v[j].i, rv.v[j].i)};
        gp[q] = np;
      }
    }
  } else {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites;
         j += (int64_t)gridDim.x * blockDim.x) {
      const int64_t hf = j / stride, i = j - hf * stride, base = hf * Sto<S>::NCOL * stride;
      cx<T> pv[12], xv[12], rv[12];
      load_site<S>(pv, p + base, i);
      load_site<S>(xv, psi + base, i);
      load_site<S>(rv, r + base, i);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        xv[k] = {fma(a, pv[k].r, xv[k].r), fma(a, pv[k].i, xv[k].i)};
        rv[k] = {fma(b, pv[k].r, rv[k].r), fma(b, pv[k].i, rv[k].i)};
      }
      store_site<S>(psi + base, xv, i, false);
      if (!keep_p) store_site<S>(p + base, rv, i, false);
    }
  }
}

template <typename S>
int launch_cg_dot(Ctx *ctx, int64_t nsites, int64_t stride, const void *p, const void *mmp, CGState *state) {
  const int nb = red_blocks(nsites);
  hipLaunchKernelGGL((k_cg_dot<S>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, (const cx<S> *)p,
                     (const cx<S> *)mmp, (const CGState *)state, ctx->red_partial);
  hipLaunchKernelGGL(k_cg_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb, state, 0);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}
template <typename S>
int launch_cg_axpy_norm(Ctx *ctx, int64_t nsites, int64_t stride, void *r, const void *mmp, CGState *state) {
  const int nb = red_blocks(nsites);
  hipLaunchKernelGGL((k_cg_axpy_norm<S>), dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, (cx<S> *)r,
                     (const cx<S> *)mmp, (const CGState *)state, ctx->red_partial);
  hipLaunchKernelGGL(k_cg_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb, state, 1);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}
template <typename S>
int launch_cg_update(Ctx *ctx, int64_t nsites, int64_t stride, void *psi, void *p, const void *r, CGState *state) {
  const int64_t work = Sto<S>::H ? nsites : 12 * nsites / Sto<S>::VW;
  hipLaunchKernelGGL((k_cg_update<S>), dim3(stream_blocks(work)), dim3(256), 0, ctx->stream, nsites, stride,
                     (cx<S> *)psi, (cx<S> *)p, (const cx<S> *)r, (const CGState *)state);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- multi-RHS slice reductions (BlockConjugateGradient.h:412-545, CGmultiRHS) ----------
// Right-hand side s of a half field is the entries e = h*Ls + s.  Each block uses floor(256/Ls)*Ls
// threads: thread t works on slice s = t % Ls and strides over the sites h, summing the 12
// components, so the per-slice partial sums come out of a fixed-order LDS reduction; a one-block
// kernel adds the block partials in block order.
__device__ __forceinline__ bool mgated(const MrhsState *st) { return gated(st); }
__device__ __forceinline__ bool mgated_prev(const MrhsState *st) {
  return st->k - 1 > st->k_conv || st->k - 1 > st->kmax;
}

template <typename S, int MODE>
__global__ void __launch_bounds__(256) k_mrhs_reduce(int64_t H4, int Ls, cx<S> *__restrict__ x,
                                                     const cx<S> *__restrict__ y, cx<S> *__restrict__ z,
                                                     const MrhsState *st, double *partial) {
  typedef typename Sto<S>::C T;
  if ((MODE == 0 || MODE == 2 || MODE == 4) && mgated(st)) return;
  __shared__ double sh[256];
  const int rpb = 256 / Ls;
  const int s = threadIdx.x % Ls;
  const int rt = threadIdx.x / Ls;
  double acc = 0;
  if (rt < rpb) {
    const T a = (MODE == 2) ? (T)(-st->a[s]) : (T)0;
    for (int64_t h = (int64_t)blockIdx.x * rpb + rt; h < H4; h += (int64_t)gridDim.x * rpb) {
      const int64_t e = h * Ls + s;
      cx<T> xv[12], yv[12];
      load_site<S>(xv, x, e);
      if (MODE != 1 && MODE != 4) load_site<S>(yv, y, e);
#pragma unroll
      for (int k = 0; k < 12; k++) {
        if (MODE == 0) {
          acc += (double)fma(xv[k].r, yv[k].r, xv[k].i * yv[k].i);
        } else if (MODE == 1 || MODE == 4) {
          acc += (double)fma(xv[k].r, xv[k].r, xv[k].i * xv[k].i);
        } else if (MODE == 2) {
          xv[k] = {fma(a, yv[k].r, xv[k].r), fma(a, yv[k].i, xv[k].i)};
          acc += (double)fma(xv[k].r, xv[k].r, xv[k].i * xv[k].i);
        } else {
          xv[k] = {yv[k].r - xv[k].r, yv[k].i - xv[k].i};
          acc += (double)fma(xv[k].r, xv[k].r, xv[k].i * xv[k].i);
        }
      }
      if (MODE == 2) store_site<S>(x, xv, e, false);
      if (MODE == 3) store_site<S>(z, xv, e, false);
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < Ls) {
    double v = 0;
    for (int r = 0; r < rpb; r++) v += sh[r * Ls + threadIdx.x];
    partial[blockIdx.x * Ls + threadIdx.x] = v;
  }
}

// which: 0, 4 -> d_s, a_s;  2 -> cp_s, b_s, convergence, k += 1;  1, 3 -> out[s] = sum
// 256 threads: thread t adds the partials b = t/Ls, t/Ls + 256/Ls, ... of slice t % Ls, and the
// per-thread sums of each slice are then added in thread order (fixed order, so deterministic).
static __global__ void __launch_bounds__(256) k_mrhs_fin(const double *partial, int nparts, int Ls, MrhsState *st,
                                                  double *out, int which) {
  if (which == 4) which = 0;
  const bool skip = (which == 0 || which == 2) && mgated(st);
  if (skip && which == 0) return;
  __shared__ double sh[256];
  const int per = 256 / Ls;
  const int ts = threadIdx.x % Ls, tp = threadIdx.x / Ls;
  double acc = 0;
  if (tp < per && !skip)
    for (int b = tp; b < nparts; b += per) acc += partial[b * Ls + ts];
  sh[threadIdx.x] = acc;
  __syncthreads();
  const int s = threadIdx.x;
  double v = 0;
  if (s < Ls)
    for (int j = 0; j < per; j++) v += sh[j * Ls + s];
  __syncthreads();
  if (which == 1 || which == 3) {
    if (s < Ls) out[s] = v;
    return;
  }
  __shared__ double rel[kMaxRhs];
  if (which == 0) {
    if (s < Ls) {
      st->d[s] = v;
      st->a[s] = (v != 0.0) ? st->c[s] / v : 0.0;
    }
    return;
  }
  if (!skip && s < Ls) {
    const double c = st->c[s];
    st->b[s] = (c != 0.0) ? v / c : 0.0;
    st->cp[s] = v;
    st->c[s] = v;
    rel[s] = v / st->ssq[s];
  }
  __syncthreads();
  if (s == 0) {
    if (!skip) {
      double mx = 0;
      for (int j = 0; j < Ls; j++) mx = rel[j] > mx ? rel[j] : mx;
      if (mx < st->tol2 && st->k_conv == INT_MAX) st->k_conv = st->k;
    }
    st->k = st->k + 1;
  }
}

// psi_s += a_s p_s; p_s = b_s p_s + r_s, one (h, s) entry (its 12 components, 12 / VW 16-byte
// groups per field, consecutive lanes on consecutive entries) per thread and step; padding entries
// stay zero.  FP16 stores rescale each entry.
template <typename T>
__global__ void __launch_bounds__(256) k_mrhs_update(int64_t nent, int Ls, cx<T> *__restrict__ psi, cx<T> *__restrict__ p,
                                                     const cx<T> *__restrict__ r, const MrhsState *st) {
  if (mgated_prev(st)) return;
  typedef typename Sto<T>::C C;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nent; e += (int64_t)gridDim.x * blockDim.x) {
    const int s = (int)(e % Ls);
    const C a = (C)st->a[s], b = (C)st->b[s];
    cx<C> pv[12], xv[12], rv[12];
    load_site<T>(pv, p, e);
    load_site<T>(xv, psi, e);
    load_site<T>(rv, r, e);
#pragma unroll
    for (int k = 0; k < 12; k++) {
      xv[k] = {fma(a, pv[k].r, xv[k].r), fma(a, pv[k].i, xv[k].i)};
      pv[k] = {fma(b, pv[k].r, rv[k].r), fma(b, pv[k].i, rv[k].i)};
    }
    store_site<T>(psi, xv, e, false);
    store_site<T>(p, pv, e, false);
  }
}

template <typename T>
int launch_mrhs_reduce(Ctx *ctx, int mode, const Geom &g, void *x, const void *y, void *z, MrhsState *state,
                       double *out_dev) {
  const int rpb = 256 / g.Ls;
  int64_t nb = (g.H4 + rpb - 1) / rpb;
  if (nb > kRedBlocks) nb = kRedBlocks;
  cx<T> *px = (cx<T> *)x;
  const cx<T> *py = (const cx<T> *)y;
  cx<T> *pz = (cx<T> *)z;
#define GX_MR(M)                                                                                                  \
  hipLaunchKernelGGL((k_mrhs_reduce<T, M>), dim3(nb), dim3(256), 0, ctx->stream, g.H4, g.Ls, px, py, pz,           \
                     (const MrhsState *)state, ctx->red_partial)
  switch (mode) {
    case 0: GX_MR(0); break;
    case 1: GX_MR(1); break;
    case 2: GX_MR(2); break;
    case 3: GX_MR(3); break;
    case 4: GX_MR(4); break;
    default: set_error("mrhs reduce mode %d", mode); return GX_EINVAL;
  }
#undef GX_MR
  hipLaunchKernelGGL(k_mrhs_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, (int)nb, g.Ls,
                     state, out_dev, mode);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

template <typename T>
int launch_mrhs_update(Ctx *ctx, const Geom &g, void *psi, void *p, const void *r, MrhsState *state) {
  const int64_t nent = g.H4 * g.Ls;  // every (h, s) entry of the half field, padding excluded
  hipLaunchKernelGGL((k_mrhs_update<T>), dim3(stream_blocks(nent)), dim3(256), 0, ctx->stream, nent, g.Ls,
                     (cx<T> *)psi, (cx<T> *)p, (const cx<T> *)r, (const MrhsState *)state);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- fifth-dimension slice transfer (LinalgUtils.h:140-183, axpby_ssp_pplus/pminus) ------
// z[s = sz] (=|+=) (spins in mask) x[s = sx] over nhalf checkerboard halves (2: full fields, 1: half
// fields); x has Lsx slices (Px padded entries per half), z has Lsz (Pz).  Used for the physical
// 4D <-> 5D source / solution maps, ExtractSlice / InsertSlice and the packing of sources into a
// multi-RHS field.
template <typename T>
__global__ void __launch_bounds__(256) k_slice_proj(const cx<T> *__restrict__ x, int Lsx, int sx, int64_t Px,
                                                    cx<T> *__restrict__ z, int Lsz, int sz, int64_t Pz, int mask,
                                                    int accumulate, int64_t H4, int nhalf) {
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nhalf * H4;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cb = j / H4, h = j - cb * H4;
    const cx<T> *xb = x + cb * 12 * Px;
    cx<T> *zb = z + cb * 12 * Pz;
    const int64_t ex = h * Lsx + sx, ez = h * Lsz + sz;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      const bool on = (mask >> (k / 3)) & 1;
      cx<T> v = on ? xb[bidx<T>(ex, k)] : cx<T>{0, 0};
      if (accumulate) {
        const cx<T> o = zb[bidx<T>(ez, k)];
        v = {o.r + v.r, o.i + v.i};
      }
      zb[bidx<T>(ez, k)] = v;
    }
  }
}

template <typename T>
int launch_slice_proj(const Geom &gx_, const void *x, int Lsx, int sx, void *z, int Lsz, int sz, int mask,
                      int accumulate, hipStream_t st, int nhalf) {
  const int64_t Px = (gx_.H4 * Lsx + kBlk - 1) / kBlk * kBlk, Pz = (gx_.H4 * Lsz + kBlk - 1) / kBlk * kBlk;
  hipLaunchKernelGGL((k_slice_proj<T>), dim3(stream_blocks(nhalf * gx_.H4)), dim3(256), 0, st, (const cx<T> *)x, Lsx,
                     sx, Px, (cx<T> *)z, Lsz, sz, Pz, mask, accumulate, gx_.H4, nhalf);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// ---------------- synthetic sources and field digests ----------------
__device__ __forceinline__ uint64_t d_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

__device__ __forceinline__ double d_u01(uint64_t key, uint64_t ctr) {
  const uint64_t r = d_splitmix64(key ^ d_splitmix64(ctr));
  return ((double)(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Element (cb, h, s0 + s, k) of a full field, s < ns, gets the Box-Muller pair of counter
// (lex * ns + s) * 12 + k, lex the lexicographic 4D site: the numbering of the host generator
// gx_gaussian_fermion for a field of ns slices, so the two agree up to the rounding of log, sqrt,
// sin and cos, and slices s0 .. s0 + ns - 1 of a wider field hold the ns-slice field of the seed.
template <typename T>
__global__ void __launch_bounds__(256) k_gaussian(cx<T> *__restrict__ out, KGeom g, int Ls, int s0, int ns,
                                                  uint64_t key) {
  const int64_t per_cb = g.H4 * ns;  // generated entries per checkerboard; g.stride = the padded count
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < 2 * per_cb;
       j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cb = j / per_cb, rel = j - cb * per_cb;
    const int64_t h = rel / ns;
    const int s = (int)(rel - h * ns);
    const int64_t rem = h * Ls + s0 + s;  // entry of the stored field
    int64_t r = h;
    const int xh = (int)(r % g.Lxh);
    r /= g.Lxh;
    const int y = (int)(r % g.L[1]);
    r /= g.L[1];
    const int z = (int)(r % g.L[2]);
    const int t = (int)(r / g.L[2]);
    const int x = 2 * xh + (int)((cb + y + z + t) & 1);
    const int64_t lex = x + (int64_t)g.L[0] * (y + (int64_t)g.L[1] * (z + (int64_t)g.L[2] * t));
    const uint64_t c0 = (uint64_t)(lex * ns + s) * 12;
    cx<T> *o = out + cb * 12 * g.stride;
#pragma unroll 2
    for (int k = 0; k < 12; k++) {
      const double a = d_u01(key, 2 * (c0 + k)), b = d_u01(key, 2 * (c0 + k) + 1);
      const double rad = sqrt(-2.0 * log(a));
      double sn, cs;
      sincos(2.0 * M_PI * b, &sn, &cs);
      o[bidx<T>(rem, k)] = {(T)(rad * cs), (T)(rad * sn)};
    }
  }
}

template <typename T>
int launch_gaussian(const Geom &g, void *full, uint64_t key, hipStream_t st, int s0, int ns) {
  if (ns <= 0) ns = g.Ls;
  if (s0 < 0 || s0 + ns > g.Ls) {
    set_error("gaussian: slices %d..%d outside Ls = %d", s0, s0 + ns - 1, g.Ls);
    return GX_EINVAL;
  }
  const KGeom k = kgeom(g);
  hipLaunchKernelGGL((k_gaussian<T>), dim3(stream_blocks(2 * g.H4 * ns)), dim3(256), 0, st, (cx<T> *)full, k, g.Ls,
                     s0, ns, key);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// digest of slices s0 .. s0 + ns - 1 of a full FP64 field over its logical elements: word i =
// (((cb * H4 + h) * ns + s - s0) * 12 + k) * 2 + part, so the slices of a wider field and the
// ns-slice field holding the same values have the same digest
static __global__ void __launch_bounds__(256) k_digest_slices(const cx<double> *__restrict__ f, int64_t H4,
                                                              int64_t stride, int Ls, int s0, int ns,
                                                              unsigned long long *out) {
  uint64_t acc = 0;
  const int64_t n = 2 * H4 * ns;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t cb = j / (H4 * ns), rel = j - cb * H4 * ns, h = rel / ns;
    const int s = (int)(rel - h * ns);
    const cx<double> *half = f + cb * 12 * stride;
    const int64_t e = h * Ls + s0 + s;
    for (int k = 0; k < 12; k++) {
      const cx<double> v = half[bidx<double>(e, k)];
      const uint64_t i = (uint64_t)(j * 12 + k) * 2;
      acc += d_splitmix64(__double_as_longlong(v.r) ^ d_splitmix64(i));
      acc += d_splitmix64(__double_as_longlong(v.i) ^ d_splitmix64(i + 1));
    }
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ uint64_t part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

// digest = sum_i splitmix64(word_i ^ splitmix64(i)) mod 2^64: integer addition is exact and
// commutative, so the result is independent of the order in which blocks finish.
static __global__ void __launch_bounds__(256) k_digest(const uint64_t *__restrict__ w, int64_t n, unsigned long long *out) {
  uint64_t acc = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    acc += d_splitmix64(w[i] ^ d_splitmix64((uint64_t)i));
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  __shared__ uint64_t part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) atomicAdd(out, (unsigned long long)(part[0] + part[1] + part[2] + part[3]));
}

// ---------------- single-direction hop and force legs (WilsonKernels DhopDirKernel,
// WilsonKernelsImplementation.h:375-410; WilsonImpl::InsertForce5D, WilsonImpl.h:193-238) ----------
// One thread per output half-site h of checkerboard cb_out, all Ls slices: the leg of direction MU
// at the forward (FWD: point MU, link Uds[MU](x)) or backward (point MU + 4, link Uds[MU + 4](x))
// neighbour with projector (1 + gamma_MU) (PLUS) or (1 - gamma_MU):
//   chi(h, s) = Uds[P](h) recon(project(in(nb_P(h), s))).
// MODE 0 writes out(h, s) = chi (DhopDir); MODE 1 writes mat[h][a][b] = sum_{s, spin} chi_a conj(A_b)
// (the colour outer product of InsertForce5D).  FP64.
template <int MU, bool FWD, bool PLUS, int MODE>
static __global__ void __launch_bounds__(256) k_hop_leg(KGeom g, int Ls, int cb_out, const cx<double> *__restrict__ in,
                                                        const cx<double> *__restrict__ U,
                                                        const cx<double> *__restrict__ A, cx<double> *__restrict__ out,
                                                        cx<double> *__restrict__ mat) {
  const uint32_t H4 = (uint32_t)g.H4;
  for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < H4; h += gridDim.x * blockDim.x) {
    const uint32_t r1 = fdiv(h, g.fLxh), r2 = fdiv(r1, g.fL1), r3 = fdiv(r2, g.fL2);
    const int xh = (int)(h - r1 * (uint32_t)g.Lxh);
    const int y = (int)(r1 - r2 * (uint32_t)g.L[1]);
    const int z = (int)(r2 - r3 * (uint32_t)g.L[2]);
    const int t = (int)r3;
    const int x = 2 * xh + ((cb_out + y + z + t) & 1);
    uint32_t n;
    if (MU == 0) {
      const int xn = FWD ? (x + 1 == g.L[0] ? 0 : x + 1) : (x == 0 ? g.L[0] - 1 : x - 1);
      n = h - xh + (xn >> 1);
    } else {
      const int c = MU == 1 ? y : (MU == 2 ? z : t), L = g.L[MU];
      const uint32_t st = (uint32_t)(MU == 1 ? g.sY : (MU == 2 ? g.sZ : g.sT));
      if (FWD) n = (c + 1 == L) ? h - (uint32_t)(L - 1) * st : h + st;
      else n = (c == 0) ? h + (uint32_t)(L - 1) * st : h - st;
    }
    constexpr int P = FWD ? MU : MU + 4;
    cx<double> u[9];
#pragma unroll
    for (int k = 0; k < 9; k++) u[k] = U[(int64_t)(P * 9 + k) * H4 + h];
    cx<double> m[9];
#pragma unroll
    for (int k = 0; k < 9; k++) m[k] = {0, 0};
    for (int s = 0; s < Ls; s++) {
      cx<double> f[12], hs[6], acc[12];
      load_site<double>(f, in, (int64_t)n * Ls + s);
      project<double, MU, PLUS>(f, hs);
#pragma unroll
      for (int k = 0; k < 12; k++) acc[k] = {0, 0};
#pragma unroll
      for (int r = 0; r < 3; r++)
        recon_colour<double, MU, PLUS>(acc, r, su3_row<double>(u + 3 * r, hs), su3_row<double>(u + 3 * r, hs + 3));
      if (MODE == 0) {
        store_site<double>(out, acc, (int64_t)h * Ls + s, false);
      } else {
        cx<double> a[12];
        load_site<double>(a, A, (int64_t)h * Ls + s);
#pragma unroll
        for (int sp = 0; sp < 4; sp++)
#pragma unroll
          for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
              const cx<double> c = acc[sp * 3 + i], d = a[sp * 3 + j];  // c conj(d)
              m[i * 3 + j].r = fma(c.r, d.r, fma(c.i, d.i, m[i * 3 + j].r));
              m[i * 3 + j].i = fma(c.i, d.r, fma(-c.r, d.i, m[i * 3 + j].i));
            }
      }
    }
    if (MODE == 1)
#pragma unroll
      for (int k = 0; k < 9; k++) mat[(int64_t)h * 9 + k] = m[k];
  }
}

// ---------------- gauge statistics (WilsonLoops.h:102-126,203-219) ----------------
// U: [site][mu][a][b] complex double, sites lexicographic (x fastest).  Per site: the real traces
// of the six plaquettes U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag (mu > nu) and of the four
// links; workgroup partials in partial[2 * block + {0, 1}], summed in a fixed order by k_reduce_fin.
struct C3 {
  double2 m[3][3];
};
static __device__ __forceinline__ C3 ld_link(const double2 *U, int64_t site, int mu) {
  C3 c;
  const double2 *p = U + (site * 4 + mu) * 9;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) c.m[a][b] = p[a * 3 + b];
  return c;
}
// x y (dx = 0) or x y^dag (dx = 1)
static __device__ __forceinline__ C3 mul3(const C3 &x, const C3 &y, bool ydag) {
  C3 z;
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      double re = 0, im = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const double2 u = x.m[a][k];
        double2 v = ydag ? y.m[b][k] : y.m[k][b];
        if (ydag) v.y = -v.y;
        re = fma(u.x, v.x, fma(-u.y, v.y, re));
        im = fma(u.x, v.y, fma(u.y, v.x, im));
      }
      z.m[a][b] = {re, im};
    }
  return z;
}
static __global__ void __launch_bounds__(256) k_gauge_stats(const double2 *__restrict__ U, int L0, int L1, int L2,
                                                            int L3, double *partial) {
  const int64_t V = (int64_t)L0 * L1 * L2 * L3;
  const int L[4] = {L0, L1, L2, L3};
  double plaq = 0, tr = 0;
  for (int64_t site = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; site < V; site += (int64_t)gridDim.x * blockDim.x) {
    int x[4];
    int64_t rem = site;
    for (int d = 0; d < 4; d++) {
      x[d] = (int)(rem % L[d]);
      rem /= L[d];
    }
    int64_t stride[4] = {1, L0, (int64_t)L0 * L1, (int64_t)L0 * L1 * L2};
    int64_t fwd[4];
    for (int d = 0; d < 4; d++) fwd[d] = site + (x[d] + 1 == L[d] ? -(int64_t)(L[d] - 1) : 1) * stride[d];
    for (int mu = 0; mu < 4; mu++) {
      const C3 umu = ld_link(U, site, mu);
      tr += umu.m[0][0].x + umu.m[1][1].x + umu.m[2][2].x;
      for (int nu = 0; nu < mu; nu++) {
        const C3 a = mul3(umu, ld_link(U, fwd[mu], nu), false);          // U_mu(x) U_nu(x+mu)
        const C3 b = mul3(ld_link(U, site, nu), ld_link(U, fwd[nu], mu), false);  // U_nu(x) U_mu(x+nu)
        // Re tr(a b^dag) = sum_ij Re(a_ij conj(b_ij))
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
          for (int j = 0; j < 3; j++) plaq += a.m[i][j].x * b.m[i][j].x + a.m[i][j].y * b.m[i][j].y;
      }
    }
  }
  __shared__ double sh[2][256];
  sh[0][threadIdx.x] = plaq;
  sh[1][threadIdx.x] = tr;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + off];
      sh[1][threadIdx.x] += sh[1][threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sh[0][0];
    partial[2 * blockIdx.x + 1] = sh[1][0];
  }
}

// ---------------- multishift CG vector updates (ConjugateGradientMultiShift.h:236-300) ------------
// One pass over r for every shift (the fusion the reference's comments propose): shift s with bit s
// of do_psi set takes psi_s -= bs_s ps_s (closing its iteration k); with bit s of do_ps set, then
// ps_s = z_s r + as_s ps_s (opening iteration k + 1); with do_p, p = a p + r.
static __global__ void __launch_bounds__(256) k_multishift(int64_t n, const cx<double> *__restrict__ r,
                                                           cx<double> *__restrict__ p, MultiShiftArgs m) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cx<double> rv = r[i];
    if (m.do_p) {
      const cx<double> pv = p[i];
      p[i] = {fma(m.a, pv.r, rv.r), fma(m.a, pv.i, rv.i)};
    }
    for (int s = 0; s < m.nshift; s++) {
      const bool dpsi = (m.do_psi >> s) & 1u, dps = (m.do_ps >> s) & 1u;
      if (!dpsi && !dps) continue;
      cx<double> *ps = (cx<double> *)m.ps[s];
      cx<double> pv = ps[i];
      if (dpsi) {
        cx<double> *psi = (cx<double> *)m.psi[s];
        const cx<double> xv = psi[i];
        psi[i] = {fma(-m.bs[s], pv.r, xv.r), fma(-m.bs[s], pv.i, xv.i)};
      }
      if (dps) {
        pv = {fma(m.z[s], rv.r, m.as[s] * pv.r), fma(m.z[s], rv.i, m.as[s] * pv.i)};
        ps[i] = pv;
      }
    }
  }
}

// ---------------- device-resident multishift recurrences ----------------
// Re<p|A p> and |p|^2 (A p before the m0 shift), block partials
static __global__ void __launch_bounds__(256) k_ms_dot(int64_t nsites, int64_t stride, const cx<double> *__restrict__ p,
                                                       const cx<double> *__restrict__ mmp, const MsState *st,
                                                       double *partial) {
  if (gated(st)) return;
  __shared__ double sh[2][4];
  double v[2] = {0, 0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride, base = hf * 12 * stride;
    cx<double> xv[12], yv[12];
    load_site<double>(xv, p + base, i);
    load_site<double>(yv, mmp + base, i);
    double sr = 0, sp = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      sr = fma(xv[k].r, yv[k].r, sr);
      sr = fma(xv[k].i, yv[k].i, sr);
      sp = fma(xv[k].r, xv[k].r, sp);
      sp = fma(xv[k].i, xv[k].i, sp);
    }
    v[0] += sr;
    v[1] += sp;
  }
  block_sum<2>(v, sh);
  write_partial<2>(v, partial);
}

// d = Re<p|A p> + m0 |p|^2, b = -cp / d, and the shifts' z and bs (ConjugateGradientMultiShift.h
// :238-262), with the a of the previous iteration
static __global__ void __launch_bounds__(256) k_ms_fin_d(const double *partial, int nparts, MsState *st) {
  if (gated(st)) return;
  __shared__ double sh[2][4];
  __shared__ double nb[2];  // the new b and the previous one
  double v[2];
  sum_partials<2>(v, partial, nparts, sh);
  if (threadIdx.x == 0) {
    const double d = v[0] + st->m0 * v[1], cp = st->c, bp = st->b, b = -cp / d;
    st->d = d;
    st->cp = cp;
    st->bp = bp;
    st->b = b;
    st->bs[0] = b;
    nb[0] = b;
    nb[1] = bp;
  }
  __syncthreads();
  const int s = threadIdx.x;
  if (s >= 1 && s < st->nshift && ((st->active >> s) & 1u)) {
    const double a = st->a, b = nb[0], bp = nb[1];
    const double z0 = st->zc[s], z1 = st->zp[s];
    const double zn = z0 * z1 * bp / (b * a * (z1 - z0) + z1 * bp * (1.0 - st->dm[s] * b));
    st->zp[s] = z0;
    st->zc[s] = zn;
    st->bs[s] = b * zn / z0;
  }
}

// r += b mmp (mmp shifted by m0) and |r|^2, block partials
static __global__ void __launch_bounds__(256) k_ms_axpy_norm(int64_t nsites, int64_t stride, cx<double> *__restrict__ r,
                                                             const cx<double> *__restrict__ mmp, const MsState *st,
                                                             double *partial) {
  if (gated(st)) return;
  __shared__ double sh[1][4];
  const double b = st->b;
  double v[1] = {0};
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < nsites; j += (int64_t)gridDim.x * blockDim.x) {
    const int64_t hf = j / stride, i = j - hf * stride, base = hf * 12 * stride;
    cx<double> xv[12], yv[12];
    load_site<double>(xv, mmp + base, i);
    load_site<double>(yv, r + base, i);
    double sr = 0;
#pragma unroll
    for (int k = 0; k < 12; k++) {
      yv[k] = {fma(b, xv[k].r, yv[k].r), fma(b, xv[k].i, yv[k].i)};
      sr = fma(yv[k].r, yv[k].r, sr);
      sr = fma(yv[k].i, yv[k].i, sr);
    }
    store_site<double>(r + base, yv, i, false);
    v[0] += sr;
  }
  block_sum<1>(v, sh);
  write_partial<1>(v, partial);
}

// c = |r|^2; each active shift's stopping test c z_s^2 < rsq_s; a = c / cp; k += 1 (also once
// stopped, so the gate of the update that closes the last iteration opens exactly once)
static __global__ void __launch_bounds__(256) k_ms_fin_c(const double *partial, int nparts, MsState *st) {
  const bool skip = gated(st);
  __shared__ double sh[1][4];
  double v[1] = {0};
  if (!skip) sum_partials<1>(v, partial, nparts, sh);
  if (threadIdx.x == 0) {
    if (!skip) {
      const double c = v[0];
      st->c = c;
      st->closing = st->active;
      uint32_t active = st->active;
      for (int s = 0; s < st->nshift; s++) {
        if (!((active >> s) & 1u)) continue;
        st->its[s] = st->k;
        if (c * st->zc[s] * st->zc[s] < st->rsq[s]) active &= ~(1u << s);
      }
      st->active = active;
      st->a = c / st->cp;
      if (active == 0 && st->k_conv == INT_MAX) st->k_conv = st->k;
    }
    st->k = st->k + 1;
  }
}

// k_multishift with the coefficients of the state (psi_s -= bs_s ps_s for the closing shifts;
// p = a p + r and ps_s = z_s r + as_s ps_s for the active ones), gated on the iteration it closes
static __global__ void __launch_bounds__(256) k_ms_update(int64_t n, const cx<double> *__restrict__ r,
                                                          cx<double> *__restrict__ p, MsPtrs m, const MsState *st) {
  if (st->k - 1 > st->k_conv || st->k - 1 > st->kmax) return;
  __shared__ double cbs[kMaxShift], cz[kMaxShift], cas[kMaxShift];
  const int ns = st->nshift;
  const uint32_t closing = st->closing, active = st->active;
  const double a = st->a;
  if (threadIdx.x < ns) {
    const int s = threadIdx.x;
    cbs[s] = st->bs[s];
    cz[s] = s == 0 ? 1.0 : st->zc[s];
    cas[s] = s == 0 ? a : a * st->zc[s] * st->bs[s] / (st->zp[s] * st->b);
  }
  __syncthreads();
  const bool do_p = active != 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const cx<double> rv = r[i];
    if (do_p) {
      const cx<double> pv = p[i];
      p[i] = {fma(a, pv.r, rv.r), fma(a, pv.i, rv.i)};
    }
    for (int s = 0; s < ns; s++) {
      const bool dpsi = (closing >> s) & 1u, dps = (active >> s) & 1u;
      if (!dpsi && !dps) continue;
      cx<double> *ps = (cx<double> *)m.ps[s];
      cx<double> pv = ps[i];
      if (dpsi) {
        cx<double> *psi = (cx<double> *)m.psi[s];
        const cx<double> xv = psi[i];
        psi[i] = {fma(-cbs[s], pv.r, xv.r), fma(-cbs[s], pv.i, xv.i)};
      }
      if (dps) {
        pv = {fma(cz[s], rv.r, cas[s] * pv.r), fma(cz[s], rv.i, cas[s] * pv.i)};
        ps[i] = pv;
      }
    }
  }
}

// ---------------- slice matrices (Lattice_reduction.h:717-776) ----------------
// FP64 fields viewed as (rows) x Ls with the fifth-dimension slice s innermost: entry e = h Ls + s
// of half hf stores component k at hf * hstride + ((e / kBlk) * 12 + k) * kBlk + e % kBlk.
// Global site g in [0, nhalf H4) is 4D half-site g % H4 of half g / H4.
static __device__ __forceinline__ int64_t slice_idx(int64_t gs, int s, int k, int64_t H4, int Ls, int64_t hstride) {
  const int64_t hf = gs >= H4 ? 1 : 0, e = (gs - hf * H4) * Ls + s;
  return hf * hstride + ((e / kBlk) * 12 + k) * kBlk + (e & (kBlk - 1));
}

// sliceInnerProductMatrix: m(s, t) = sum conj(x_s) y_t.  A workgroup stages kIpmSites sites (every
// slice and component) of x and y in LDS; thread s * Ls + t accumulates m(s, t) over them and
// writes its workgroup's partial sum; k_slice_ipm_fin adds the partials in a fixed order.
constexpr int kIpmSites = 8;
static __global__ void __launch_bounds__(256) k_slice_ipm(int64_t nsite, int64_t H4, int Ls, int64_t hstride,
                                                          const cx<double> *__restrict__ x,
                                                          const cx<double> *__restrict__ y, double2 *partial) {
  __shared__ double2 sx[kIpmSites * kMaxLs * 12], sy[kIpmSites * kMaxLs * 12];
  const int t = threadIdx.x, ls2 = Ls * Ls;
  const int si = t / Ls, ti = t - si * Ls;
  const int per = kIpmSites * Ls;
  double ar = 0, ai = 0;
  for (int64_t g0 = (int64_t)blockIdx.x * kIpmSites; g0 < nsite; g0 += (int64_t)gridDim.x * kIpmSites) {
    const int ns = (int)min((int64_t)kIpmSites, nsite - g0);
    __syncthreads();
    for (int q = t; q < 12 * per; q += 256) {
      const int k = q / per, j = q - k * per;  // component k, staged entry j = site * Ls + s
      double2 vx = {0, 0}, vy = {0, 0};
      if (j < ns * Ls) {
        const int hs = j / Ls;
        const int64_t i = slice_idx(g0 + hs, j - hs * Ls, k, H4, Ls, hstride);
        vx = {x[i].r, x[i].i};
        vy = {y[i].r, y[i].i};
      }
      sx[j * 12 + k] = vx;
      sy[j * 12 + k] = vy;
    }
    __syncthreads();
    if (t < ls2)
      for (int hs = 0; hs < ns; hs++) {
        const double2 *px = sx + (hs * Ls + si) * 12, *py = sy + (hs * Ls + ti) * 12;
#pragma unroll
        for (int k = 0; k < 12; k++) {
          ar = fma(px[k].x, py[k].x, fma(px[k].y, py[k].y, ar));
          ai = fma(px[k].x, py[k].y, fma(-px[k].y, py[k].x, ai));
        }
      }
  }
  if (t < ls2) partial[(int64_t)blockIdx.x * ls2 + t] = {ar, ai};
}

// workgroup p adds the partials of matrix entry p: thread t the partials t, t + 256, ... in order,
// then a fixed tree over the threads (bitwise reproducible)
static __global__ void __launch_bounds__(256) k_slice_ipm_fin(const double2 *partial, int nparts, int ls2,
                                                              double2 *out) {
  const int t = threadIdx.x, p = blockIdx.x;
  double ar = 0, ai = 0;
  for (int b = t; b < nparts; b += 256) {
    ar += partial[(int64_t)b * ls2 + p].x;
    ai += partial[(int64_t)b * ls2 + p].y;
  }
  __shared__ double sh[2][256];
  sh[0][t] = ar;
  sh[1][t] = ai;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (t < off) {
      sh[0][t] += sh[0][t + off];
      sh[1][t] += sh[1][t + off];
    }
    __syncthreads();
  }
  if (t == 0) out[p] = {sh[0][0], sh[1][0]};
}

// sliceMaddMatrix: r_t = y_t + scale sum_s x_s a(s, t), a row-major (a[s * Ls + t]); y = null reads
// as zero (sliceMulMatrix).  A workgroup stages kMaddSites sites of x (every slice and component)
// in LDS with coalesced loads, then each thread forms outputs (site, t, component) from the staged
// slices and writes them, again coalesced.  Every x of a stage is read before any r of it is
// written (and y at an index only by the thread that writes r there), so r may alias x or y.
constexpr int kMaddSites = 4;
static __global__ void __launch_bounds__(256) k_slice_madd(int64_t nsite, int64_t H4, int Ls, int64_t hstride,
                                                           const double2 *a, double scale, const cx<double> *x,
                                                           const cx<double> *y, cx<double> *r) {
  __shared__ double2 sa[kMaxLs * kMaxLs];
  __shared__ double2 sx[kMaddSites * kMaxLs * 12];
  const int t = threadIdx.x;
  for (int q = t; q < Ls * Ls; q += blockDim.x) sa[q] = {scale * a[q].x, scale * a[q].y};
  const int per = kMaddSites * Ls;
  for (int64_t g0 = (int64_t)blockIdx.x * kMaddSites; g0 < nsite; g0 += (int64_t)gridDim.x * kMaddSites) {
    const int ns = (int)min((int64_t)kMaddSites, nsite - g0);
    __syncthreads();  // the previous stage's LDS reads are done (and sa is written)
    for (int q = t; q < 12 * per; q += 256) {
      const int k = q / per, j = q - k * per;  // component k, staged entry j = site * Ls + s
      if (j < ns * Ls) {
        const int hs = j / Ls;
        const cx<double> v = x[slice_idx(g0 + hs, j - hs * Ls, k, H4, Ls, hstride)];
        sx[j * 12 + k] = {v.r, v.i};
      }
    }
    __syncthreads();
    for (int q = t; q < 12 * per; q += 256) {
      const int k = q / per, j = q - k * per;  // output component k, entry j = site * Ls + tt
      if (j < ns * Ls) {
        const int hs = j / Ls, tt = j - hs * Ls;
        const int64_t idx = slice_idx(g0 + hs, tt, k, H4, Ls, hstride);
        double orr = 0, oi = 0;
        if (y) {
          const cx<double> v = y[idx];
          orr = v.r, oi = v.i;
        }
        const double2 *xs = sx + hs * Ls * 12 + k;
        for (int s2 = 0; s2 < Ls; s2++) {
          const double2 xv = xs[s2 * 12], c = sa[s2 * Ls + tt];
          orr = fma(xv.x, c.x, fma(-xv.y, c.y, orr));
          oi = fma(xv.x, c.y, fma(xv.y, c.x, oi));
        }
        r[idx] = {orr, oi};
      }
    }
  }
}

// ---------------- matrices over lists of fields (BlockConjugateGradient.h:35-74) ----------------
// InnerProductMatrix: m(b, bp) = <X_b, Y_bp> over n FP64 fields of nelem stored complex numbers
// each (padding entries are zero).  A workgroup stages kVecChunk elements of every X_b and Y_b in
// LDS; thread b * n + bp accumulates m(b, bp) over them and writes its workgroup's partial sum,
// which k_slice_ipm_fin adds in a fixed order.
constexpr int kVecChunk = 64;
static __global__ void __launch_bounds__(256) k_vec_ipm(int64_t nelem, int n, VecPtrs x, VecPtrs y, double2 *partial) {
  __shared__ double2 sx[kMaxLs * kVecChunk], sy[kMaxLs * kVecChunk];
  const int t = threadIdx.x, n2 = n * n;
  const int bi = t / n, bj = t - bi * n;
  double ar = 0, ai = 0;
  for (int64_t i0 = (int64_t)blockIdx.x * kVecChunk; i0 < nelem; i0 += (int64_t)gridDim.x * kVecChunk) {
    const int ne = (int)min((int64_t)kVecChunk, nelem - i0);
    __syncthreads();
    for (int q = t; q < n * kVecChunk; q += 256) {
      const int f = q / kVecChunk, j = q - f * kVecChunk;  // field f, element i0 + j (coalesced in j)
      double2 vx = {0, 0}, vy = {0, 0};
      if (j < ne) {
        vx = reinterpret_cast<const double2 *>(x.p[f])[i0 + j];
        vy = reinterpret_cast<const double2 *>(y.p[f])[i0 + j];
      }
      sx[q] = vx;
      sy[q] = vy;
    }
    __syncthreads();
    if (t < n2) {
      const double2 *px = sx + bi * kVecChunk, *py = sy + bj * kVecChunk;
      for (int j = 0; j < ne; j++) {
        ar = fma(px[j].x, py[j].x, fma(px[j].y, py[j].y, ar));
        ai = fma(px[j].x, py[j].y, fma(-px[j].y, py[j].x, ai));
      }
    }
  }
  if (t < n2) partial[(int64_t)blockIdx.x * n2 + t] = {ar, ai};
}

// MaddMatrix: R_b = Y_b + scale sum_bp X_bp a(bp, b) (Y null reads as zero: MulMatrix), one thread per
// element with every X_bp of that element in registers before any R_b is written, so R may alias X
// or Y as in Grid.
static __global__ void __launch_bounds__(256) k_vec_madd(int64_t nelem, int n, const double2 *a, double scale, VecPtrs x,
                                                         VecPtrs y, VecPtrs r) {
  __shared__ double2 sa[kMaxLs * kMaxLs];
  for (int q = threadIdx.x; q < n * n; q += blockDim.x) sa[q] = {scale * a[q].x, scale * a[q].y};
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nelem; i += (int64_t)gridDim.x * blockDim.x) {
    double2 xv[kMaxLs];
#pragma unroll
    for (int b = 0; b < kMaxLs; b++)
      if (b < n) xv[b] = reinterpret_cast<const double2 *>(x.p[b])[i];
    for (int b = 0; b < n; b++) {
      double2 o = y.p[b] ? reinterpret_cast<const double2 *>(y.p[b])[i] : double2{0, 0};
#pragma unroll
      for (int bp = 0; bp < kMaxLs; bp++)
        if (bp < n) {
          const double2 c = sa[bp * n + b];
          o.x = fma(xv[bp].x, c.x, fma(-xv[bp].y, c.y, o.x));
          o.y = fma(xv[bp].x, c.y, fma(xv[bp].y, c.x, o.y));
        }
      reinterpret_cast<double2 *>(const_cast<void *>(r.p[b]))[i] = o;
    }
  }
}

#if GX_PART == 0
// the per-tile slice sums of a norm-producing hop folded into kRedBlocks block partials: thread t
// of a block adds the tiles blockIdx * rpb + t / Ls, + gridDim * rpb, ... of slice t % Ls, the
// threads of a slice are then added in thread order (fixed order, so deterministic)
static __global__ void __launch_bounds__(256) k_fold_tiles(const double *__restrict__ tp, int64_t ntiles, int Ls,
                                                           const MrhsState *st, double *partial) {
  if (mgated(st)) return;
  __shared__ double sh[256];
  const int rpb = 256 / Ls;
  const int s = threadIdx.x % Ls, rt = threadIdx.x / Ls;
  double acc = 0;
  if (rt < rpb)
    for (int64_t t = (int64_t)blockIdx.x * rpb + rt; t < ntiles; t += (int64_t)gridDim.x * rpb) acc += tp[t * Ls + s];
  sh[threadIdx.x] = acc;
  __syncthreads();
  if (threadIdx.x < Ls) {
    double v = 0;
    for (int r = 0; r < rpb; r++) v += sh[r * Ls + threadIdx.x];
    partial[blockIdx.x * Ls + threadIdx.x] = v;
  }
}

int launch_fold_norms(Ctx *ctx, const double *npart, int64_t ntiles, int Ls, MrhsState *state, int which) {
  if (Ls < 1 || Ls > kMaxRhs || ntiles < 1 || (which != 2 && which != 4)) {
    set_error("fold norms: %lld tiles of %d slices (which %d)", (long long)ntiles, Ls, which);
    return GX_EINVAL;
  }
  const int rpb = 256 / Ls;
  int64_t nb = (ntiles + rpb - 1) / rpb;
  if (nb > kRedBlocks) nb = kRedBlocks;
  hipLaunchKernelGGL(k_fold_tiles, dim3(nb), dim3(256), 0, ctx->stream, npart, ntiles, Ls, (const MrhsState *)state,
                     ctx->red_partial);
  hipLaunchKernelGGL(k_mrhs_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, (int)nb, Ls,
                     state, nullptr, which);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

// the same tile sums added over every tile and slice into block partials of the single-RHS CG
// (fixed order: thread, wave, workgroup)
static __global__ void __launch_bounds__(256) k_fold_all(const double *__restrict__ tp, int64_t n, const CGState *st,
                                                         double *partial) {
  if (gated(st)) return;
  __shared__ double sh[1][4];
  double v[1] = {0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[0] += tp[i];
  block_sum<1>(v, sh);
  write_partial<1>(v, partial);
}

int launch_fold_cg(Ctx *ctx, const double *npart, int64_t ntiles, int Ls, CGState *state, int which) {
  if (Ls < 1 || ntiles < 1 || (which != 0 && which != 1)) {
    set_error("fold cg: %lld tiles of %d slices (which %d)", (long long)ntiles, Ls, which);
    return GX_EINVAL;
  }
  const int64_t n = ntiles * Ls;
  const int nb = red_blocks(n);
  hipLaunchKernelGGL(k_fold_all, dim3(nb), dim3(256), 0, ctx->stream, npart, n, (const CGState *)state,
                     ctx->red_partial);
  hipLaunchKernelGGL(k_cg_fin, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb, state, which);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_vec_ipm(Ctx *ctx, int64_t nelem, int n, const VecPtrs &x, const VecPtrs &y, void *partial, int nparts,
                   void *out_dev) {
  if (n < 1 || n > kMaxLs || nparts < 1) {
    set_error("inner product matrix: %d fields (1..%d)", n, kMaxLs);
    return GX_EINVAL;
  }
  int64_t nb = (nelem + kVecChunk - 1) / kVecChunk;
  if (nb > nparts) nb = nparts;
  hipLaunchKernelGGL(k_vec_ipm, dim3(nb), dim3(256), 0, ctx->stream, nelem, n, x, y, (double2 *)partial);
  hipLaunchKernelGGL(k_slice_ipm_fin, dim3(n * n), dim3(256), 0, ctx->stream, (const double2 *)partial, (int)nb, n * n,
                     (double2 *)out_dev);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_vec_madd(Ctx *ctx, int64_t nelem, int n, const void *a_dev, double scale, const VecPtrs &x, const VecPtrs &y,
                    const VecPtrs &r) {
  if (n < 1 || n > kMaxLs) {
    set_error("madd matrix: %d fields (1..%d)", n, kMaxLs);
    return GX_EINVAL;
  }
  hipLaunchKernelGGL(k_vec_madd, dim3(stream_blocks(nelem)), dim3(256), 0, ctx->stream, nelem, n, (const double2 *)a_dev,
                     scale, x, y, r);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_slice_ipm(Ctx *ctx, const Geom &g, int nhalf, const void *x, const void *y, void *partial, int nparts,
                     void *out_dev) {
  if (g.Ls > kMaxLs || nparts < 1) {
    set_error("slice inner product matrix: Ls %d > %d or no partials", g.Ls, kMaxLs);
    return GX_EINVAL;
  }
  const int64_t nsite = nhalf * g.H4, hstride = 12 * g.half_pad();
  int64_t nb = (nsite + kIpmSites - 1) / kIpmSites;
  if (nb > nparts) nb = nparts;
  hipLaunchKernelGGL(k_slice_ipm, dim3(nb), dim3(256), 0, ctx->stream, nsite, g.H4, g.Ls, hstride,
                     (const cx<double> *)x, (const cx<double> *)y, (double2 *)partial);
  hipLaunchKernelGGL(k_slice_ipm_fin, dim3(g.Ls * g.Ls), dim3(256), 0, ctx->stream, (const double2 *)partial,
                     (int)nb, g.Ls * g.Ls, (double2 *)out_dev);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_hop_leg(const Geom &g, int mu, int fwd, int plus, int mode, int cb_out, const void *in, const void *U,
                   const void *A, void *out, void *mat, hipStream_t st) {
  if (mu < 0 || mu > 3) {
    set_error("hop leg: direction %d", mu);
    return GX_EINVAL;
  }
  const KGeom k = kgeom(g);
  const uint32_t nb = (uint32_t)((g.H4 + 255) / 256);
  const cx<double> *pin = (const cx<double> *)in, *pU = (const cx<double> *)U, *pA = (const cx<double> *)A;
  cx<double> *pout = (cx<double> *)out, *pmat = (cx<double> *)mat;
#define GX_LEG(M, F, P, MODE)                                                                                    \
  hipLaunchKernelGGL((k_hop_leg<M, F, P, MODE>), dim3(nb), dim3(256), 0, st, k, g.Ls, cb_out, pin, pU, pA, pout, pmat)
#define GX_LEG_MODE(M, F, P) \
  if (mode == 0) GX_LEG(M, F, P, 0); else GX_LEG(M, F, P, 1)
#define GX_LEG_FP(M)                                              \
  if (fwd) { if (plus) GX_LEG_MODE(M, true, true); else GX_LEG_MODE(M, true, false); } \
  else { if (plus) GX_LEG_MODE(M, false, true); else GX_LEG_MODE(M, false, false); }
  switch (mu) {
    case 0: GX_LEG_FP(0); break;
    case 1: GX_LEG_FP(1); break;
    case 2: GX_LEG_FP(2); break;
    default: GX_LEG_FP(3); break;
  }
#undef GX_LEG_FP
#undef GX_LEG_MODE
#undef GX_LEG
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_gauge_stats(Ctx *ctx, const int dims[4], const void *U_dev, double *partial_dev, int nparts,
                       double *out_host) {
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  int64_t nb = (V + 255) / 256;
  if (nb > nparts) nb = nparts;
  hipLaunchKernelGGL(k_gauge_stats, dim3(nb), dim3(256), 0, ctx->stream, (const double2 *)U_dev, dims[0], dims[1],
                     dims[2], dims[3], partial_dev);
  GX_CHECK_HIP(hipGetLastError());
  GX_CHECK_HIP(hipMemcpyAsync(out_host, partial_dev, sizeof(double) * 2 * nb, hipMemcpyDeviceToHost, ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  double p = 0, t = 0;
  for (int64_t b = 0; b < nb; b++) {  // fixed order: bitwise reproducible
    p += out_host[2 * b];
    t += out_host[2 * b + 1];
  }
  out_host[0] = p;
  out_host[1] = t;
  return GX_OK;
}

int launch_ms_dot(Ctx *ctx, int64_t nsites, int64_t stride, const void *p, const void *mmp, MsState *st) {
  const int nb = red_blocks(nsites);
  hipLaunchKernelGGL(k_ms_dot, dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, (const cx<double> *)p,
                     (const cx<double> *)mmp, (const MsState *)st, ctx->red_partial);
  hipLaunchKernelGGL(k_ms_fin_d, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb, st);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_ms_axpy_norm(Ctx *ctx, int64_t nsites, int64_t stride, void *r, const void *mmp, MsState *st) {
  const int nb = red_blocks(nsites);
  hipLaunchKernelGGL(k_ms_axpy_norm, dim3(nb), dim3(256), 0, ctx->stream, nsites, stride, (cx<double> *)r,
                     (const cx<double> *)mmp, (const MsState *)st, ctx->red_partial);
  hipLaunchKernelGGL(k_ms_fin_c, dim3(1), dim3(256), 0, ctx->stream, (const double *)ctx->red_partial, nb, st);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_ms_update(Ctx *ctx, int64_t n, const void *r, void *p, const MsPtrs &m, MsState *st) {
  hipLaunchKernelGGL(k_ms_update, dim3(stream_blocks(n)), dim3(256), 0, ctx->stream, n, (const cx<double> *)r,
                     (cx<double> *)p, m, (const MsState *)st);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_multishift(Ctx *ctx, int64_t n, const void *r, void *p, const MultiShiftArgs &m) {
  if (m.nshift < 1 || m.nshift > kMaxShift) {
    set_error("multishift update: %d shifts (1..%d)", m.nshift, kMaxShift);
    return GX_EINVAL;
  }
  hipLaunchKernelGGL(k_multishift, dim3(stream_blocks(n)), dim3(256), 0, ctx->stream, n, (const cx<double> *)r,
                     (cx<double> *)p, m);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_slice_madd(Ctx *ctx, const Geom &g, int nhalf, const void *a_dev, double scale, const void *x,
                      const void *y, void *r) {
  if (g.Ls > kMaxLs) {
    set_error("slice madd matrix: Ls %d > %d", g.Ls, kMaxLs);
    return GX_EINVAL;
  }
  const int64_t nsite = nhalf * g.H4, hstride = 12 * g.half_pad();
  int64_t nb = (nsite + kMaddSites - 1) / kMaddSites;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_slice_madd, dim3(nb), dim3(256), 0, ctx->stream, nsite, g.H4, g.Ls,
                     hstride, (const double2 *)a_dev, scale, (const cx<double> *)x, (const cx<double> *)y,
                     (cx<double> *)r);
  GX_CHECK_HIP(hipGetLastError());
  return GX_OK;
}

int launch_digest(Ctx *ctx, const void *data, int64_t nwords, uint64_t *out) {
  GX_CHECK_HIP(hipMemsetAsync(ctx->digest, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_digest, dim3(stream_blocks(nwords)), dim3(256), 0, ctx->stream, (const uint64_t *)data, nwords,
                     ctx->digest);
  GX_CHECK_HIP(hipGetLastError());
  GX_CHECK_HIP(hipMemcpyAsync(ctx->pinned, ctx->digest, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out, ctx->pinned, sizeof(uint64_t));
  return GX_OK;
}

int launch_digest_slices(Ctx *ctx, const Geom &g, const void *full, int s0, int ns, uint64_t *out) {
  if (s0 < 0 || ns < 1 || s0 + ns > g.Ls) {
    set_error("digest: slices %d..%d outside Ls = %d", s0, s0 + ns - 1, g.Ls);
    return GX_EINVAL;
  }
  GX_CHECK_HIP(hipMemsetAsync(ctx->digest, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_digest_slices, dim3(stream_blocks(2 * g.H4 * ns)), dim3(256), 0, ctx->stream,
                     (const cx<double> *)full, g.H4, g.half_pad(), g.Ls, s0, ns, ctx->digest);
  GX_CHECK_HIP(hipGetLastError());
  GX_CHECK_HIP(hipMemcpyAsync(ctx->pinned, ctx->digest, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                              ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  std::memcpy(out, ctx->pinned, sizeof(uint64_t));
  return GX_OK;
}
#endif

#define GX_INST(S)                                                                                                 \
  template int launch_dhop_epi<S>(const Geom &, const void *, void *, const void *, int, int, const Gate *,       \
                                  hipStream_t, int, const void *, const void *, double, double, int, const void *, \
                                  const HopFuse *);                                                                \
  template int launch_smat2<S>(const Geom &, const void *, const void *, double, void *, const void *, void *,     \
                               const void *, const Gate *, hipStream_t, int);                                    \
  template int launch_dhop<S>(const Geom &, const void *, void *, const void *, int, int, const Gate *,           \
                              hipStream_t);                                                                        \
  template int launch_smat<S>(const Geom &, const void *, const void *, void *, const void *, const Gate *,        \
                              hipStream_t, int);                                                                   \
  template int launch_axpby<S>(int64_t, int64_t, double, const void *, double, const void *, void *, const Gate *, \
                               hipStream_t);                                                                       \
  template int launch_scale_slices<S>(const Geom &, const void *, void *, const void *, const Gate *, hipStream_t); \
  template int launch_reduce<S>(Ctx *, int, int64_t, int64_t, const void *, const void *, void *, double, int);   \
  template int launch_cg_dot<S>(Ctx *, int64_t, int64_t, const void *, const void *, CGState *);                   \
  template int launch_cg_axpy_norm<S>(Ctx *, int64_t, int64_t, void *, const void *, CGState *);                   \
  template int launch_cg_update<S>(Ctx *, int64_t, int64_t, void *, void *, const void *, CGState *);
#define GX_INST_REAL(T)                                                                                            \
  template int launch_slice_proj<T>(const Geom &, const void *, int, int, void *, int, int, int, int, hipStream_t, \
                                    int);                                                                          \
  template int launch_mrhs_reduce<T>(Ctx *, int, const Geom &, void *, const void *, void *, MrhsState *, double *); \
  template int launch_mrhs_update<T>(Ctx *, const Geom &, void *, void *, const void *, MrhsState *);             \
  template int launch_gaussian<T>(const Geom &, void *, uint64_t, hipStream_t, int, int);
#if GX_PART == 0
GX_INST(double)
GX_INST_REAL(double)
template int launch_convert<double, float>(int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<double, float>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<float, double>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_accum_site<float>(int64_t, int64_t, const void *, void *, hipStream_t);
#elif GX_PART == 1
GX_INST(float)
GX_INST_REAL(float)
#else
GX_INST(half_t)
template int launch_mrhs_reduce<half_t>(Ctx *, int, const Geom &, void *, const void *, void *, MrhsState *, double *);
template int launch_mrhs_update<half_t>(Ctx *, const Geom &, void *, void *, const void *, MrhsState *);
template int launch_convert<double, half_t>(int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<double, half_t>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<half_t, double>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<float, half_t>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_convert_site<half_t, float>(int64_t, int64_t, const void *, void *, hipStream_t);
template int launch_accum_site<half_t>(int64_t, int64_t, const void *, void *, hipStream_t);
#endif

}  // namespace gx
