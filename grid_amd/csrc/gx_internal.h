// Internal declarations of libgx: host-side objects and kernel launchers.
//
// Device layout (HBM), chosen for coalesced wave64 access on gfx950:
//   * every field is stored per checkerboard; a full field is [even half | odd half];
//   * a half field holds entries e = h*Ls + s (h = half-lattice 4D site in Grid's red-black order,
//     x halved; s = fifth-dimension / right-hand-side index, innermost) of the 12 spin-colour
//     components k = spin*3 + colour, in blocks of 64 entries: data[e/64][k][e%64].  A wave whose
//     lanes walk (s, h) with s fastest reads 64 consecutive complex numbers per component, and the
//     12 components of those entries are one contiguous run; FP16 fields add a 13th column of
//     per-site float scales;
//   * doubled gauge links per output checkerboard, SoA over sites:
//       U_cb[p][k][h],  p = 0..7 (x+,y+,z+,t+,x-,y-,z-,t-), k = a*3+b;
//     U_cb[p<4] = -1/2 U_mu(x) (boundary phase), U_cb[p>=4] = -1/2 U_mu(x-mu)^dag.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gx.h"

namespace gx {

typedef _Float16 half_t;  // FP16 storage (GX_F16)

void set_error(const char *fmt, ...);

// f(S()) with S the storage type of precision prec (double, float, half_t)
template <class F>
int by_prec(int prec, F &&f) {
  if (prec == GX_F64) return f(double());
  if (prec == GX_F32) return f(float());
  if (prec == GX_F16) return f(half_t());
  set_error("precision %d unsupported", prec);
  return GX_EINVAL;
}

// storage precision of a storage type
template <typename S>
constexpr int prec_of() {
  return sizeof(S) == 8 ? GX_F64 : (sizeof(S) == 4 ? GX_F32 : GX_F16);
}

// process-wide launch-shape knobs (gx_tune_set / gx_tune_get), one set per storage precision
// (prec < 0: the FP64 set); the version changes on every set
int tune_get(int knob, int prec = -1);
int tune_version();

#define GX_CHECK_HIP(call)                                                                             \
  do {                                                                                                 \
    hipError_t e_ = (call);                                                                            \
    if (e_ != hipSuccess) {                                                                            \
      gx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_));                 \
      return GX_EHIP;                                                                                  \
    }                                                                                                  \
  } while (0)

#define GX_TRY(call)                                                                                   \
  do {                                                                                                 \
    int r_ = (call);                                                                                   \
    if (r_ != GX_OK) return r_;                                                                        \
  } while (0)

constexpr int kMaxLs = 16;         // largest extent of the templated kernels and of the block solvers' blocks
constexpr int kMaxFifth = 128;     // largest fifth-dimension extent (Ls, or right-hand sides x Ls)
constexpr int kMaxRhs = 64;        // right-hand sides of one multi-RHS solve
constexpr int kRedBlocks = 1024;   // fixed grid of the deterministic reductions
constexpr int kRedSlots = 64;      // device result slots per context

// Fields are stored in blocks of kBlk consecutive entries (h, s) of one checkerboard; within a block
// the 12 components of every entry are stored in groups of vw (the complex numbers that fill 16
// bytes: 1 for FP64, 2 for FP32, 4 for FP16), group-major.  Component k < 12 of entry e lives at
//   (e / kBlk) ncol kBlk + (k / vw) kBlk vw + (e % kBlk) vw + k % vw
// (in complex numbers of the storage type), so a wave of 64 consecutive entries moves one
// contiguous run per group and each lane 16 bytes per access.  FP16 adds column 12, one float per
// entry (k = 12: (e / kBlk) ncol kBlk + 12 kBlk + e % kBlk).
constexpr int kBlk = 64;
inline int64_t blk_index(int64_t e, int k, int ncol, int vw) {
  const int64_t base = (e / kBlk) * ncol * kBlk;
  if (k >= 12) return base + 12 * kBlk + e % kBlk;
  return base + (int64_t)(k / vw) * kBlk * vw + (e % kBlk) * vw + k % vw;
}

struct Geom {
  int L[4];
  int Ls;
  int64_t V4, H4;
  int64_t half_elems() const { return H4 * Ls; }   // entries (h, s) per checkerboard
  int64_t half_pad() const { return (H4 * Ls + kBlk - 1) / kBlk * kBlk; }  // entries incl. block padding
};

// Iteration gate shared by the device-resident solvers: a kernel of iteration k runs only while
// k <= min(k_conv, kmax), so a host can enqueue iterations ahead and read the outcome later.
struct Gate {
  int k;        // current iteration (1-based)
  int k_conv;   // iteration at which the stopping test first passed (INT_MAX until then)
  int kmax;     // last iteration allowed (MaxIterations)
  int pad;
};

// Device-resident scalars of one CG (ConjugateGradient.h:68-257).
// With delta > 0 (reliable-update CG, ConjugateGradientReliableUpdate.h:140-252) the norm
// finaliser also tracks the largest residual since the last reliable update and, when the
// residual falls below delta times it, stops the iteration with ru = 1 so the host can recompute
// the residual in FP64; the solution update of that iteration then leaves p for the host.
struct CGState : Gate {
  double c, cp, a, b, d, rsq;
  double delta, maxres, cprev;  // cprev: c of the last finished iteration
  double fbtol;                 // > 0: stop when cp < fbtol so the host can switch precision
  int ru, pad2;                 // ru: 1 reliable update pending, 2 precision switch pending
};

// Per-RHS solver scalars for the multi-RHS CG (right-hand side = slice s of the fifth dimension).
// As in Grid's CGmultiRHS every right-hand side iterates until the largest relative residual
// max_s cp_s / ssq_s drops below tol^2; k and k_conv gate the kernels as in CGState.
struct MrhsState : Gate {
  double c[kMaxRhs], cp[kMaxRhs], a[kMaxRhs], b[kMaxRhs], d[kMaxRhs], ssq[kMaxRhs];
  double tol2;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev[16] = {};
  double *red_partial = nullptr;     // [kRedBlocks * kMaxRhs]
  unsigned long long *digest = nullptr;  // field digest accumulator
  double *red_result = nullptr;      // [kRedSlots]
  double *pinned = nullptr;          // host mirror of red_result
  int next_slot = 0;
};

struct Field {
  Ctx *ctx = nullptr;
  Geom g{};
  int prec = GX_F64;
  int subset = GX_FULL;   // GX_FULL or a half field
  int cb = 0;             // checkerboard tag of a half field
  // half: ncol() columns (component k = column k) over g.half_pad() entries in the block layout
  // (blk_index); full: 2 halves (even, odd).  GX_F16 stores a complex number as two binary16 values
  // and adds a 13th column: one float per (h, s) site, the site's largest |component|, by which
  // its stored values are multiplied.  Padding entries hold zeros.
  void *data = nullptr;
  size_t bytes = 0;
  int halves() const { return subset == GX_FULL ? 2 : 1; }
  int64_t elems() const { return 12 * g.half_elems() * halves(); }           // logical complex numbers
  int64_t stored() const { return 12 * g.half_pad() * halves(); }            // stored, padding included
  size_t csize() const { return prec == GX_F64 ? 16 : (prec == GX_F32 ? 8 : 4); }
  int ncol() const { return prec == GX_F16 ? 13 : 12; }
  size_t half_bytes() const { return (size_t)ncol() * g.half_pad() * csize(); }
  // pointer to the half with checkerboard cb (full fields) or the data (half fields)
  void *half(int which) const {
    if (subset != GX_FULL) return data;
    return (char *)data + (size_t)which * half_bytes();
  }
};

// Per-chirality Ls x Ls real matrices of a fifth-dimension operator, row-major:
// mat[0] acts on spins 0,1 (P+), mat[1] on spins 2,3 (P-).
struct SMat {
  std::vector<double> m;   // 2*Ls*Ls
};

// Device state, pinned mirrors, events and captured graphs of the solvers of one operator
// (gx_solver.hip); released with the operator.
struct SolverWork;
void solver_work_free(SolverWork *w);
// device bytes of the solver state that scale with the lattice (the multi-RHS tile sums)
size_t solver_work_bytes(const SolverWork *w);

struct Op {
  Ctx *ctx = nullptr;
  int kind = GX_WILSON;
  Geom g{};
  double mass = 0, M5 = 0, b = 0, c = 0;
  double phases[8] = {1, 0, 1, 0, 1, 0, 1, 0};
  void *U[3][2] = {};           // [prec][output cb] doubled links
  std::vector<double> host_U;   // host copy of the gauge field (lexicographic), for re-import
  // Cayley coefficients (CayleyFermion5DImplementation.h:410-535)
  std::vector<double> bs, cs, bee, cee, dee, lee, leem, uee, ueem;
  // fifth-dimension operators as per-chirality matrices: 0 Meooe5D, 1 MeooeDag5D, 2 Mooee,
  // 3 MooeeDag, 4 MooeeInv, 5 MooeeInvDag, 6 M5D off-diagonal part, 7 M5Ddag off-diagonal part,
  // 8 Meooe5D * MooeeInv, 9 MooeeInvDag * MeooeDag5D (products used by the fused Schur operator)
  enum {
    S_MEOOE5D = 0, S_MEOOEDAG5D, S_MOOEE, S_MOOEEDAG, S_MOOEEINV, S_MOOEEINVDAG, S_M5D_OFF, S_M5DDAG_OFF,
    S_MEO5D_INV, S_INVDAG_MEODAG5D, S_COUNT
  };
  SMat smat[S_COUNT];
  void *smat_dev[3] = {};       // [prec] device copy of all S_COUNT matrices (2*Ls*Ls each)
  // per precision temporaries (half fields): 0-3 operator compositions, 4-6 CG vectors,
  // 7-8 mixed-precision outer vectors, 10-14 the red-black solve's source and solution halves
  static constexpr int kTmp = 16;
  Field *tmp[3][kTmp] = {};
  Field *full_tmp = nullptr;    // F64 full field: the red-black solve's residual check
  SolverWork *work = nullptr;   // solver state, created on the first solve
  uint64_t epoch = 0;           // bumped whenever links, mass or matrices change (invalidates graphs)
  // Per-slice coefficients of the Schur operator, in FP64 (coef_dev[0]) and FP32 (coef_dev[1], also
  // FP16 storage), 4 Ls values: [0, Ls) alpha_s and [Ls, 2 Ls) beta_s of the fused Mpc epilogue
  // (Wilson -1/d_s, d_s with d_s = 4 + m_s; Cayley -1, 1), [2 Ls, 3 Ls) 1/d_s, [3 Ls, 4 Ls) d_s.
  // slice_mass: Wilson per-slice masses m_s (empty: every slice has `mass`).
  std::vector<double> slice_mass;
  void *coef_dev[2] = {};
  std::vector<double> coef_host;   // host sources of the coef_dev copies
  std::vector<float> coef_host_f;
  Op *link_owner = nullptr;     // clones read this operator's links and own none
  bool owns_links = true;       // false for clones: the links belong to link_owner
  int clones = 0;               // live clones reading this operator's links (guarded by the share lock)
};

// ---------------- launchers (gx_kernels.hip) ----------------
// full field of Gaussian complex numbers with the host generator's counter numbering; with ns > 0
// only slices s0 .. s0 + ns - 1, holding the ns-slice field of the key
template <typename T>
int launch_gaussian(const Geom &g, void *full, uint64_t key, hipStream_t st, int s0 = 0, int ns = 0);
// order-independent 64-bit digest of nwords 64-bit words: sum over i of mix(word_i, i) mod 2^64
int launch_digest(Ctx *ctx, const void *data, int64_t nwords, uint64_t *out);
// the same over the logical elements of slices s0 .. s0 + ns - 1 of a full FP64 field
int launch_digest_slices(Ctx *ctx, const Geom &g, const void *full, int s0, int ns, uint64_t *out);
// All pointers are device pointers to fields in the layout above.  S is the storage type: double,
// float or half_t (FP16 storage with float arithmetic; its matrices and scalars are float).
template <typename S>
int launch_dhop(const Geom &g, const void *in, void *out, const void *U, int cb_out, int dag, const Gate *gate,
                hipStream_t st);
// hop with an epilogue: epi 0 out = hop; 1 out_s = alpha_s hop_s + beta_s phi_s with per-slice
// coefficients coef (device, 2 Ls in T: alpha_s then beta_s); 2 out = alpha A(hop) [+ beta phi] with A
// a per-chirality matrix applied across each site's s-row, block diagonal in groups of lb consecutive
// slices (2*lb*lb in T; lb = 0: one group of Ls).  fz: reductions fused into the epilogue (HopFuse).
//
// HopFuse, on Ls >= 2 where the tuned kernel is the 4-wave whole-site one or a general shape:
//   r == nullptr, epi 1, dag 0: out as epi 1, and the per-tile slice sums of |out|^2 in npart;
//   r != nullptr, epi 1 or 2 with phi, dag 1: r_s -= a[astride s] (the epilogue's result)_s and out
//   is NOT written; per-tile slice sums of |r_s|^2 in npart (a on the device: astride 1, one value
//   per slice (multi-RHS); astride 0, one value (single-RHS CG)).
// npart holds [tile][Ls] doubles, at most hop_norm_tiles(g.H4, g.Ls) tiles; *nparts = the tiles
// written, or 0 when the kernel has no such epilogue: out is then written as plain epi 1 and the
// caller reduces itself.
struct HopFuse {
  double *npart = nullptr;
  int64_t *nparts = nullptr;
  void *r = nullptr;
  const double *a = nullptr;
  int astride = 1;
};
template <typename T>
int launch_dhop_epi(const Geom &g, const void *in, void *out, const void *U, int cb_out, int dag, const Gate *gate,
                    hipStream_t st, int epi, const void *phi, const void *mat, double alpha, double beta, int lb = 0,
                    const void *coef = nullptr, const HopFuse *fz = nullptr);
// upper bound of the tiles of a norm-producing hop over H4 half sites of Ls slices (the 4-wave
// kernel has 4 (64 / Ls) sites per tile, a general shape of W waves 64 W / Ls >= 64 / Ls), and the
// bytes of their slice sums
inline int64_t hop_norm_tiles(int64_t H4, int Ls) {
  const int spb = Ls <= 64 ? 64 / Ls : 1;
  return (H4 + spb - 1) / spb;
}
size_t hop_norm_bytes(int64_t H4, int Ls);
// per-tile slice sums left by a fused hop (ntiles tiles of Ls) folded, in fixed order, into the
// multi-RHS CG state as the reduction of that mode: which 4 -> d_s, a_s; 2 -> cp_s, b_s,
// convergence, k += 1
int launch_fold_norms(Ctx *ctx, const double *npart, int64_t ntiles, int Ls, MrhsState *state, int which);
// the same tile sums added over all slices into the single-RHS CG state: which 0 -> d, a; 1 -> cp, b,
// stopping test, k += 1 (as the dot and axpy-norm finalisers)
int launch_fold_cg(Ctx *ctx, const double *npart, int64_t ntiles, int Ls, CGState *state, int which);
// out_e = c[s] in_e on the (h, s) entries of half fields (c: Ls values in the arithmetic type)
template <typename S>
int launch_scale_slices(const Geom &g, const void *in, void *out, const void *c, const Gate *gate, hipStream_t st);
// out = A_chi(in) per chirality (+ phi if phi), mats = 2*lb*lb device matrix in T (groups as above)
template <typename T>
int launch_smat(const Geom &g, const void *in, const void *phi, void *out, const void *mats, const Gate *gate,
                hipStream_t st, int lb = 0);
// out = A(in) + beta phi and, if out2, out2 = C(in) from one read of in
template <typename T>
int launch_smat2(const Geom &g, const void *in, const void *phi, double beta, void *out, const void *mats, void *out2,
                 const void *mats2, const Gate *gate, hipStream_t st, int lb = 0);
// z = a x + b y over nsites sites (stride sites per half; a, b real; y optional)
template <typename S>
int launch_axpby(int64_t nsites, int64_t stride, double a, const void *x, double b, const void *y, void *z,
                 const Gate *gate, hipStream_t st);
// reductions into ctx slot(s); mode 0: <x|y> -> slot, slot+1 (re, im); 1: |x|^2 -> slot;
// 2: z = a x + y and |z|^2 -> slot (fused axpy_norm); 3: Re<x|y> -> slot, |x|^2 -> slot+1
template <typename T>
int launch_reduce(Ctx *ctx, int mode, int64_t nsites, int64_t stride, const void *x, const void *y, void *z, double a,
                  int slot);
template <typename Tin, typename Tout>
int launch_convert(int64_t n, const void *in, void *out, hipStream_t st);       // n plain entries
template <typename Sin, typename Sout>
int launch_convert_site(int64_t nsites, int64_t stride, const void *in, void *out, hipStream_t st);
// out (FP64) += in (FP32 or FP16 storage), site by site
template <typename Sin>
int launch_accum_site(int64_t nsites, int64_t stride, const void *in, void *out, hipStream_t st);
// z[sz] (=|+=) (spins in mask) x[sx] over nhalf checkerboard halves (2: full fields, 1: half fields;
// rows Lsx / Lsz)
template <typename T>
int launch_slice_proj(const Geom &g, const void *x, int Lsx, int sx, void *z, int Lsz, int sz, int mask,
                      int accumulate, hipStream_t st, int nhalf = 2);

// CG iteration kernels on device scalars (gx_kernels.hip)
template <typename T>
int launch_cg_dot(Ctx *ctx, int64_t nsites, int64_t stride, const void *p, const void *mmp, CGState *state);
template <typename T>
int launch_cg_axpy_norm(Ctx *ctx, int64_t nsites, int64_t stride, void *r, const void *mmp, CGState *state);
template <typename S>
int launch_cg_update(Ctx *ctx, int64_t nsites, int64_t stride, void *psi, void *p, const void *r, CGState *state);

// Multi-RHS (slice) kernels; a half field is viewed as rows x Ls with s innermost.
// mode 0: d_s = Re<x_s|y_s> -> a_s = c_s / d_s (state);   mode 4: d_s = |x_s|^2 -> a_s (state);
// mode 2: x_s -= a_s y_s, cp_s = |x_s|^2 -> b_s, convergence, k += 1 (state);   mode 1: out_s = |x_s|^2;
// mode 3: z_s = y_s - x_s, out_s = |z_s|^2
template <typename T>
int launch_mrhs_reduce(Ctx *ctx, int mode, const Geom &g, void *x, const void *y, void *z, MrhsState *state,
                       double *out_dev);
template <typename T>
int launch_mrhs_update(Ctx *ctx, const Geom &g, void *psi, void *p, const void *r, MrhsState *state);

// Multishift CG vector updates over n stored complex numbers of FP64 fields (one pass over r):
// bit s of do_psi: psi_s -= bs_s ps_s; then bit s of do_ps: ps_s = z_s r + as_s ps_s; do_p: p = a p + r.
constexpr int kMaxShift = 32;
struct MultiShiftArgs {
  void *ps[kMaxShift];
  void *psi[kMaxShift];
  double bs[kMaxShift], z[kMaxShift], as[kMaxShift];
  double a;
  uint32_t do_psi, do_ps;
  int nshift, do_p;
};
int launch_multishift(Ctx *ctx, int64_t n, const void *r, void *p, const MultiShiftArgs &m);
// Device-resident scalars of the multishift CG from its second iteration on (Grid's recurrences,
// ConjugateGradientMultiShift.h:225-300): the lightest system's c, cp, a, b, bp, d and each shift's
// z (current zc, previous zp), bs and stopping norm rsq; active: shifts still iterating (bit s),
// closing: the shifts whose psi the next update closes.  k, k_conv gate the kernels as in CGState.
struct MsState : Gate {
  double c, cp, a, b, bp, d, m0;
  double zc[kMaxShift], zp[kMaxShift], bs[kMaxShift], dm[kMaxShift], rsq[kMaxShift];
  uint32_t active, closing;
  int nshift, pad;
  int its[kMaxShift];
};
struct MsPtrs {
  void *ps[kMaxShift];
  void *psi[kMaxShift];
};
// one multishift iteration's reductions and vector updates on FP64 half fields of nsites sites
// (stride per half) with the state on the device: d from (p, mmp) (mmp = A p, before the m0
// shift) and the z / bs recurrences; r += b mmp with |r|^2, the stopping tests and a; then the
// psi / p / ps updates of k_multishift with the state's coefficients.  dot: 0 -> after the hop,
// 1 -> after mmp += m0 p
int launch_ms_dot(Ctx *ctx, int64_t nsites, int64_t stride, const void *p, const void *mmp, MsState *st);
int launch_ms_axpy_norm(Ctx *ctx, int64_t nsites, int64_t stride, void *r, const void *mmp, MsState *st);
int launch_ms_update(Ctx *ctx, int64_t n, const void *r, void *p, const MsPtrs &m, MsState *st);
// Sums over the lattice of the six plaquettes' and four links' real traces of a device gauge field
// ([site][mu][3][3] complex double): out_host[0] (plaquettes), out_host[1] (links); out_host holds
// 2 nparts doubles.  Synchronises the stream.
int launch_gauge_stats(Ctx *ctx, const int dims[4], const void *U_dev, double *partial_dev, int nparts,
                       double *out_host);
// One leg of the FP64 hop on the half fields of checkerboard cb_out (links U of cb_out): direction
// mu, forward (fwd) or backward neighbour, projector (1 + gamma) (plus) or (1 - gamma).  mode 0:
// out = the leg (half field of cb_out);  mode 1: mat[h][3][3] = sum over s and spin of leg (x) A^dag.
int launch_hop_leg(const Geom &g, int mu, int fwd, int plus, int mode, int cb_out, const void *in, const void *U,
                   const void *A, void *out, void *mat, hipStream_t st);

// Slice matrices of FP64 fields with nhalf (1 or 2) checkerboard halves (sliceInnerProductMatrix,
// sliceMaddMatrix).  ipm: out_dev[s * Ls + t] = sum conj(x_s) y_t (double2), through nparts
// workgroup partials of Ls^2 double2 each.  madd: r_t = y_t + scale sum_s x_s a(s, t), a_dev
// row-major Ls x Ls double2, y may be null (zero); r may alias x or y.
int launch_slice_ipm(Ctx *ctx, const Geom &g, int nhalf, const void *x, const void *y, void *partial, int nparts,
                     void *out_dev);
int launch_slice_madd(Ctx *ctx, const Geom &g, int nhalf, const void *a_dev, double scale, const void *x,
                      const void *y, void *r);

// Matrices over lists of n <= kMaxLs FP64 fields of nelem stored complex numbers each
// (InnerProductMatrix, MaddMatrix of BlockConjugateGradient.h:35-74).  ipm: out_dev[b * n + bp] =
// <x_b, y_bp> (double2) through nparts workgroup partials of n^2 double2 each.  madd: r_b = y_b +
// scale sum_bp x_bp a(bp, b), a_dev row-major n x n double2; y.p[b] may be null (zero); r may
// alias x or y.
struct VecPtrs {
  const void *p[kMaxLs];
};
int launch_vec_ipm(Ctx *ctx, int64_t nelem, int n, const VecPtrs &x, const VecPtrs &y, void *partial, int nparts,
                   void *out_dev);
int launch_vec_madd(Ctx *ctx, int64_t nelem, int n, const void *a_dev, double scale, const VecPtrs &x, const VecPtrs &y,
                    const VecPtrs &r);

// ---------------- api helpers (gx_api.hip) ----------------
int field_create(Ctx *ctx, const Geom &g, int prec, int subset, Field **out);
void field_destroy(Field *f);
Field *op_tmp(Op *op, int prec, int i);
int op_gauge(Op *op, int prec, void **Ue, void **Uo);
bool same_geom(const Geom &a, const Geom &b);
Field half_view(const Field *full, int cb);  // non-owning view of one checkerboard of a full field
int op_prepare(Op *op, int prec);  // allocate lazily created resources of a precision ahead of capture
int op_smat(Op *op, int prec, int which, const void **mats);
int op_coef(Op *op, int prec, const void **coef);  // coef_dev of the arithmetic precision of prec
int op_upload_coefs(Op *op);                       // (re)build coef_dev from mass / slice_mass
int reduce_read(Ctx *ctx, int slot, int n, double *out);

// Operator building blocks on half fields (in/out precision must match)
int apply_dhop_half(Op *op, Field *in, Field *out, int cb_out, int dag, const Gate *gate = nullptr);
int apply_meooe(Op *op, Field *in, Field *out, int dag, const Gate *gate = nullptr);
int apply_mooee(Op *op, Field *in, Field *out, int which, const Gate *gate = nullptr);
// fz: HopFuse for the second hop (Wilson, fused Mpc path; *fz->nparts = 0 otherwise)
int apply_mpc(Op *op, Field *in, Field *out, int dag, const Gate *gate = nullptr, const HopFuse *fz = nullptr);
// out = MpcDag Mpc in through temporary 2 (left holding Mpc in)
int apply_mpcdagmpc(Op *op, Field *in, Field *out, const Gate *gate = nullptr);
int apply_full_m(Op *op, Field *in, Field *out, int dag);
int blas_axpby(Field *z, double a, Field *x, double b, Field *y, const Gate *gate = nullptr);
int blas_norm2(Field *x, double *nrm);
int blas_inner(Field *x, Field *y, double *re, double *im);
int blas_inner_norm(Field *x, Field *y, double *re, double *nrm);  // Re<x|y>, |x|^2 in one pass
int blas_axpy_norm(Field *z, double a, Field *x, Field *y, double *nrm);
int convert_field(Field *out, Field *in);
// out += in for an FP64 out and an FP32 / FP16-storage in of one geometry, both full or both half
int accumulate_field(Field *out, Field *in);
// ExtractSlice (extract) / InsertSlice between a 4D field lo (Ls = 1) and slice s of hi: two full or
// two half fields of one precision
int slice_transfer(Field *lo, Field *hi, int s, bool extract);

}  // namespace gx

// The ABI's opaque handles are the library's objects.
struct gx_ctx_s : public gx::Ctx {};
struct gx_field_s : public gx::Field {};
struct gx_op_s : public gx::Op {};
