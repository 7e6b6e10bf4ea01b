// libgx C ABI: contexts, fields, operators, BLAS1 and the even-odd solvers.
//
// Operator compositions follow Grid (paths relative to the Grid source tree):
//   Wilson   M = Dhop + (4+m), Meooe = DhopEO/OE by input checkerboard, Mooee = (4+m),
//            MooeeInv = 1/(4+m)          WilsonFermionImplementation.h:113-176,310-342
//   Cayley   Meooe = Dhop o Meooe5D, MeooeDag = MeooeDag5D o Dhop^dag,
//            Mooee, MooeeDag, MooeeInv(Dag), M5D(dag)   CayleyFermion5DImplementation.h:155-329
//   Schur    Mpc = Mooee - Meooe MooeeInv Meooe       LinearOperator.h:325-349
//   CG       ConjugateGradient.h:68-257; mixed precision ConjugateGradientMixedPrec.h:71-165;
//            red-black solve SchurRedBlack.h:236-291,385-425.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>

#include "gx_internal.h"

namespace gx {

static thread_local char g_err[1024] = "";

// knob values per storage precision (GX_F64, GX_F32, GX_F16): a mixed-precision solve runs its
// FP64 and its inner operators each with the launch shape tuned for that precision
static int g_tune[3][GX_TUNE_COUNT] = {{1, 1, 1, 0, 1, 1, 0, 0, 0, 0},   // map pf nt order graph fused shape
                                       {1, 1, 1, 0, 1, 1, 0, 0, 0, 0},   // persist mfma tile
                                       {1, 1, 1, 0, 1, 1, 0, 0, 0, 0}};
static int g_tune_version = 0;
int tune_get(int knob, int prec) {
  if (knob < 0 || knob >= GX_TUNE_COUNT) return 0;
  return g_tune[(prec >= GX_F64 && prec <= GX_F16) ? prec : GX_F64][knob];
}
int tune_version() { return g_tune_version; }

void set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------- host parallel loop ----------------
static void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &f) {
  int nt = (int)std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  if (nt > 32) nt = 32;
  if (n < 4096 || nt == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  const int64_t chunk = (n + nt - 1) / nt;
  for (int t = 0; t < nt; t++) {
    const int64_t b = t * chunk, e = std::min<int64_t>(n, b + chunk);
    if (b >= e) break;
    th.emplace_back([=, &f]() { f(b, e); });
  }
  for (auto &x : th) x.join();
}

static bool geom_ok(const int dims[4], int Ls) {
  if (!dims) return false;
  for (int d = 0; d < 4; d++)
    if (dims[d] < 2) return false;
  if (dims[0] % 2) return false;
  return Ls >= 1 && Ls <= kMaxFifth;
}

static Geom make_geom(const int dims[4], int Ls) {
  Geom g;
  for (int d = 0; d < 4; d++) g.L[d] = dims[d];
  g.Ls = Ls;
  g.V4 = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  g.H4 = g.V4 / 2;
  return g;
}

bool same_geom(const Geom &a, const Geom &b) {
  for (int d = 0; d < 4; d++)
    if (a.L[d] != b.L[d]) return false;
  return a.Ls == b.Ls;
}

static inline void coor_lex(const Geom &g, int64_t idx, int x[4]) {
  for (int d = 0; d < 4; d++) {
    x[d] = (int)(idx % g.L[d]);
    idx /= g.L[d];
  }
}
static inline int64_t lex_index(const Geom &g, const int x[4]) {
  return x[0] + (int64_t)g.L[0] * (x[1] + (int64_t)g.L[1] * (x[2] + (int64_t)g.L[2] * x[3]));
}
static inline int64_t half_index(const Geom &g, const int x[4]) {
  return (x[0] >> 1) + (int64_t)(g.L[0] / 2) * (x[1] + (int64_t)g.L[1] * (x[2] + (int64_t)g.L[2] * x[3]));
}
static inline void coor_half(const Geom &g, int64_t h, int cb, int x[4]) {
  const int lxh = g.L[0] / 2;
  const int xh = (int)(h % lxh);
  h /= lxh;
  x[1] = (int)(h % g.L[1]);
  h /= g.L[1];
  x[2] = (int)(h % g.L[2]);
  h /= g.L[2];
  x[3] = (int)h;
  x[0] = 2 * xh + ((cb + x[1] + x[2] + x[3]) & 1);
}

// ---------------- fields ----------------
int field_create(Ctx *ctx, const Geom &g, int prec, int subset, Field **out) {
  if (prec != GX_F64 && prec != GX_F32 && prec != GX_F16) {
    set_error("field precision %d not supported", prec);
    return GX_EINVAL;
  }
  Field *f = new Field();
  f->ctx = ctx;
  f->g = g;
  f->prec = prec;
  f->subset = subset;
  f->cb = subset == GX_FULL ? 0 : subset;
  f->bytes = f->half_bytes() * (subset == GX_FULL ? 2 : 1);
  hipError_t e = hipMalloc(&f->data, f->bytes);
  if (e != hipSuccess) {
    set_error("hipMalloc(%zu) failed: %s", f->bytes, hipGetErrorString(e));
    delete f;
    return GX_ENOMEM;
  }
  e = hipMemsetAsync(f->data, 0, f->bytes, ctx->stream);
  if (e != hipSuccess) {
    (void)hipFree(f->data);
    delete f;
    GX_CHECK_HIP(e);
  }
  *out = f;
  return GX_OK;
}

void field_destroy(Field *f) {
  if (!f) return;
  if (f->data) (void)hipFree(f->data);
  delete f;
}

// A non-owning half-field view of one checkerboard of a full field.
Field half_view(const Field *full, int cb) {
  Field v = *full;
  v.data = full->half(cb);
  v.subset = cb;
  v.cb = cb;
  return v;
}

Field *op_tmp(Op *op, int prec, int i) {
  Field *&t = op->tmp[prec][i];
  if (!t) {
    if (field_create(op->ctx, op->g, prec, GX_EVEN, &t) != GX_OK) return nullptr;
  }
  return t;
}

// Creates everything the operator applications of one precision allocate on first use (its link
// copy and temporaries 0-3), so that they can be captured into a graph afterwards.
int op_prepare(Op *op, int prec) {
  void *ue, *uo;
  GX_TRY(op_gauge(op, prec, &ue, &uo));
  for (int i = 0; i < 4; i++)
    if (!op_tmp(op, prec, i)) return GX_ENOMEM;
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  return GX_OK;
}

// ---------------- reductions ----------------
int reduce_read(Ctx *ctx, int slot, int n, double *out) {
  GX_CHECK_HIP(hipMemcpyAsync(ctx->pinned + slot, ctx->red_result + slot, sizeof(double) * n, hipMemcpyDeviceToHost,
                              ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; i++) out[i] = ctx->pinned[slot + i];
  return GX_OK;
}

// entries of a field as the BLAS and reduction kernels walk them: every stored entry of each half
// (padding entries hold zeros and contribute nothing)
static int64_t nsites_of(const Field *f) { return f->g.half_pad() * f->halves(); }

static bool conformable(const Field *a, const Field *b) {
  return a && b && same_geom(a->g, b->g) && a->prec == b->prec && (a->subset == GX_FULL) == (b->subset == GX_FULL);
}

int blas_norm2(Field *x, double *nrm) {
  GX_TRY(by_prec(x->prec, [&](auto t) {
    return launch_reduce<decltype(t)>(x->ctx, 1, nsites_of(x), x->g.half_pad(), x->data, nullptr, nullptr, 0, 0);
  }));
  return reduce_read(x->ctx, 0, 1, nrm);
}

int blas_inner(Field *x, Field *y, double *re, double *im) {
  GX_TRY(by_prec(x->prec, [&](auto t) {
    return launch_reduce<decltype(t)>(x->ctx, 0, nsites_of(x), x->g.half_pad(), x->data, y->data, nullptr, 0, 0);
  }));
  double v[2];
  GX_TRY(reduce_read(x->ctx, 0, 2, v));
  *re = v[0];
  if (im) *im = v[1];
  return GX_OK;
}

// Re<x|y> and |x|^2 in one pass and one read-back
int blas_inner_norm(Field *x, Field *y, double *re, double *nrm) {
  GX_TRY(by_prec(x->prec, [&](auto t) {
    return launch_reduce<decltype(t)>(x->ctx, 3, nsites_of(x), x->g.half_pad(), x->data, y->data, nullptr, 0, 0);
  }));
  double v[2];
  GX_TRY(reduce_read(x->ctx, 0, 2, v));
  *re = v[0];
  *nrm = v[1];
  return GX_OK;
}

int blas_axpy_norm(Field *z, double a, Field *x, Field *y, double *nrm) {
  GX_TRY(by_prec(x->prec, [&](auto t) {
    return launch_reduce<decltype(t)>(x->ctx, 2, nsites_of(x), x->g.half_pad(), x->data, y->data, z->data, a, 0);
  }));
  z->cb = x->cb;
  return reduce_read(x->ctx, 0, 1, nrm);
}

int blas_axpby(Field *z, double a, Field *x, double b, Field *y, const Gate *gate) {
  z->cb = x->cb;
  return by_prec(x->prec, [&](auto t) {
    return launch_axpby<decltype(t)>(nsites_of(x), x->g.half_pad(), a, x->data, b, y ? y->data : nullptr, z->data,
                                     gate, x->ctx->stream);
  });
}

// Precisions store a site's components in different groupings (16 bytes per group), so every
// conversion between precisions works site by site.
int convert_field(Field *out, Field *in) {
  out->cb = in->cb;
  if (in->prec == out->prec) {
    GX_CHECK_HIP(hipMemcpyAsync(out->data, in->data, in->bytes, hipMemcpyDeviceToDevice, in->ctx->stream));
    return GX_OK;
  }
  const int64_t ns = nsites_of(in), st = in->g.half_pad();
  hipStream_t s = in->ctx->stream;
  if (in->prec == GX_F64 && out->prec == GX_F32) return launch_convert_site<double, float>(ns, st, in->data, out->data, s);
  if (in->prec == GX_F32 && out->prec == GX_F64) return launch_convert_site<float, double>(ns, st, in->data, out->data, s);
  if (in->prec == GX_F64 && out->prec == GX_F16) return launch_convert_site<double, half_t>(ns, st, in->data, out->data, s);
  if (in->prec == GX_F16 && out->prec == GX_F64) return launch_convert_site<half_t, double>(ns, st, in->data, out->data, s);
  if (in->prec == GX_F32 && out->prec == GX_F16) return launch_convert_site<float, half_t>(ns, st, in->data, out->data, s);
  if (in->prec == GX_F16 && out->prec == GX_F32) return launch_convert_site<half_t, float>(ns, st, in->data, out->data, s);
  set_error("precision change %d -> %d unsupported", in->prec, out->prec);
  return GX_EINVAL;
}

int accumulate_field(Field *out, Field *in) {
  if (out->prec != GX_F64 || in->prec == GX_F64 || out->halves() != in->halves() || !same_geom(out->g, in->g)) {
    set_error("accumulate: needs an F64 field and an F32 / F16 field of one lattice, both full or both half");
    return GX_EINVAL;
  }
  const int64_t ns = nsites_of(in), st = in->g.half_pad();
  hipStream_t s = in->ctx->stream;
  if (in->prec == GX_F32) return launch_accum_site<float>(ns, st, in->data, out->data, s);
  return launch_accum_site<half_t>(ns, st, in->data, out->data, s);
}

// ---------------- gauge field: DoubleStore into the device layout ----------------
// Uds[mu] = -1/2 U_mu(x) * phase (x_mu = L-1); Uds[mu+4] = -1/2 U_mu(x-mu)^dag * conj(phase) (x_mu = 0)
// WilsonImpl.h:127-171, WilsonFermionImplementation.h:82-107.
static void double_store(const Geom &g, const double *U, const double *ph, int cb, std::vector<double> &out) {
  const int64_t H4 = g.H4;
  out.assign((size_t)8 * 9 * H4 * 2, 0.0);
  parallel_for(H4, [&](int64_t b, int64_t e) {
    for (int64_t h = b; h < e; h++) {
      int x[4];
      coor_half(g, h, cb, x);
      const int64_t s = lex_index(g, x);
      for (int mu = 0; mu < 4; mu++) {
        const double pr = ph[2 * mu], pi = ph[2 * mu + 1];
        // forward
        {
          const double *u = U + (s * 4 + mu) * 18;
          const bool edge = (x[mu] == g.L[mu] - 1);
          for (int k = 0; k < 9; k++) {
            double re = u[2 * k], im = u[2 * k + 1];
            if (edge) {
              const double r2 = re * pr - im * pi, i2 = re * pi + im * pr;
              re = r2;
              im = i2;
            }
            double *o = &out[(((size_t)mu * 9 + k) * H4 + h) * 2];
            o[0] = -0.5 * re;
            o[1] = -0.5 * im;
          }
        }
        // backward: U_mu(x-mu)^dag
        {
          int y[4] = {x[0], x[1], x[2], x[3]};
          y[mu] = (x[mu] - 1 + g.L[mu]) % g.L[mu];
          const double *u = U + (lex_index(g, y) * 4 + mu) * 18;
          const bool edge = (x[mu] == 0);
          for (int a = 0; a < 3; a++)
            for (int bb = 0; bb < 3; bb++) {
              double re = u[2 * (bb * 3 + a)], im = -u[2 * (bb * 3 + a) + 1];
              if (edge) {  // * conj(phase)
                const double r2 = re * pr + im * pi, i2 = im * pr - re * pi;
                re = r2;
                im = i2;
              }
              double *o = &out[(((size_t)(mu + 4) * 9 + a * 3 + bb) * H4 + h) * 2];
              o[0] = -0.5 * re;
              o[1] = -0.5 * im;
            }
        }
      }
    }
  });
}

static void free_gauge(Op *op) {
  for (int p = 0; p < 3; p++)
    for (int cb = 0; cb < 2; cb++)
      if (op->U[p][cb]) {
        if (op->owns_links) (void)hipFree(op->U[p][cb]);
        op->U[p][cb] = nullptr;
      }
}

static int upload_gauge(Op *op) {
  free_gauge(op);
  const size_t n = (size_t)8 * 9 * op->g.H4;  // complex per checkerboard
  for (int cb = 0; cb < 2; cb++) {
    std::vector<double> h;
    double_store(op->g, op->host_U.data(), op->phases, cb, h);
    GX_CHECK_HIP(hipMalloc(&op->U[GX_F64][cb], n * 16));
    GX_CHECK_HIP(hipMemcpy(op->U[GX_F64][cb], h.data(), n * 16, hipMemcpyHostToDevice));
  }
  return GX_OK;
}

// Links of precision prec, derived from the FP64 ones on first use (FP16: unscaled binary16).
// A clone takes them from the operator that owns its links.
static std::mutex g_share_mu;
int op_gauge(Op *op, int prec, void **Ue, void **Uo) {
  if (!op->U[prec][0] && op->link_owner) {
    std::lock_guard<std::mutex> lk(g_share_mu);
    Op *own = op->link_owner;
    if (!own->U[prec][0]) {
      const size_t n = (size_t)8 * 9 * own->g.H4;
      if (prec != GX_F32 && prec != GX_F16) {
        set_error("gauge precision %d unsupported", prec);
        return GX_EINVAL;
      }
      for (int cb = 0; cb < 2; cb++) {
        GX_CHECK_HIP(hipMalloc(&own->U[prec][cb], n * (prec == GX_F32 ? 8 : 4)));
        if (prec == GX_F32)
          GX_TRY((launch_convert<double, float>(2 * n, own->U[GX_F64][cb], own->U[prec][cb], own->ctx->stream)));
        else
          GX_TRY((launch_convert<double, half_t>(2 * n, own->U[GX_F64][cb], own->U[prec][cb], own->ctx->stream)));
      }
      GX_CHECK_HIP(hipStreamSynchronize(own->ctx->stream));
    }
    op->U[prec][0] = own->U[prec][0];
    op->U[prec][1] = own->U[prec][1];
  }
  if (!op->U[prec][0]) {
    const size_t n = (size_t)8 * 9 * op->g.H4;
    if (prec != GX_F32 && prec != GX_F16) {
      set_error("gauge precision %d unsupported", prec);
      return GX_EINVAL;
    }
    for (int cb = 0; cb < 2; cb++) {
      GX_CHECK_HIP(hipMalloc(&op->U[prec][cb], n * (prec == GX_F32 ? 8 : 4)));
      if (prec == GX_F32)
        GX_TRY((launch_convert<double, float>(2 * n, op->U[GX_F64][cb], op->U[prec][cb], op->ctx->stream)));
      else
        GX_TRY((launch_convert<double, half_t>(2 * n, op->U[GX_F64][cb], op->U[prec][cb], op->ctx->stream)));
    }
  }
  *Ue = op->U[prec][0];
  *Uo = op->U[prec][1];
  return GX_OK;
}

// ---------------- Cayley coefficients and fifth-dimension matrices ----------------
// SetCoefficientsInternal, CayleyFermion5DImplementation.h:410-535, Mobius tanh (omega_s = 1).
static void cayley_coeffs(Op *op) {
  const int Ls = op->g.Ls;
  const double b = op->b, c = op->c, M5 = op->M5, mass = op->mass;
  auto &bs = op->bs, &cs = op->cs, &bee = op->bee, &cee = op->cee;
  bs.assign(Ls, 0);
  cs.assign(Ls, 0);
  bee.assign(Ls, 0);
  cee.assign(Ls, 0);
  const double bpc = b + c, bmc = b - c;
  for (int i = 0; i < Ls; i++) {
    bs[i] = 0.5 * (bpc + bmc);
    cs[i] = 0.5 * (bpc - bmc);
    bee[i] = bs[i] * (4.0 - M5) + 1.0;
    cee[i] = 1.0 - cs[i] * (4.0 - M5);
  }
  const std::vector<double> &aee = cee;
  op->dee.assign(Ls, 0);
  op->lee.assign(Ls, 0);
  op->leem.assign(Ls, 0);
  op->uee.assign(Ls, 0);
  op->ueem.assign(Ls, 0);
  for (int i = 0; i < Ls; i++) {
    op->dee[i] = bee[i];
    if (i < Ls - 1) {
      op->lee[i] = -cee[i + 1] / bee[i];
      op->leem[i] = mass * cee[Ls - 1] / bee[0];
      for (int j = 0; j < i; j++) op->leem[i] *= aee[j] / bee[j + 1];
      op->uee[i] = -aee[i] / bee[i];
      op->ueem[i] = mass;
      for (int j = 1; j <= i; j++) op->ueem[i] *= cee[j] / bee[j];
      op->ueem[i] *= aee[0] / bee[0];
    }
  }
  double delta_d = mass * cee[Ls - 1];
  for (int j = 0; j < Ls - 1; j++) delta_d *= cee[j] / bee[j];
  op->dee[Ls - 1] += delta_d;
}

// Matrices of a Grid M5D-type kernel (CayleyFermion5Dcache.h:41-115):
//   non-dag: chi_s = diag_s phi_s + upper_s P- psi_{s+1} + lower_s P+ psi_{s-1}
//   dag:     chi_s = diag_s phi_s + upper_s P+ psi_{s+1} + lower_s P- psi_{s-1}
// P+ acts on spins 0,1 (matrix 0), P- on spins 2,3 (matrix 1).
static void m5d_matrix(int Ls, const std::vector<double> &lower, const std::vector<double> &diag,
                       const std::vector<double> &upper, bool dag, SMat &out) {
  out.m.assign(2 * Ls * Ls, 0.0);
  double *Ap = out.m.data(), *Am = out.m.data() + Ls * Ls;
  for (int s = 0; s < Ls; s++) {
    const int su = (s + 1) % Ls, sl = (s + Ls - 1) % Ls;
    Ap[s * Ls + s] += diag[s];
    Am[s * Ls + s] += diag[s];
    if (!dag) {
      Am[s * Ls + su] += upper[s];
      Ap[s * Ls + sl] += lower[s];
    } else {
      Ap[s * Ls + su] += upper[s];
      Am[s * Ls + sl] += lower[s];
    }
  }
}

// MooeeInv / MooeeInvDag as dense per-chirality matrices, built column by column by running
// Grid's LDU sweeps (CayleyFermion5Dcache.h:117-230) on unit vectors in FP64.
static void mooee_inv_matrix(const Op *op, bool dag, SMat &out) {
  const int Ls = op->g.Ls;
  out.m.assign(2 * Ls * Ls, 0.0);
  for (int chi = 0; chi < 2; chi++) {
    const bool up = (chi == 0);
    const bool carry = dag ? !up : up;
    const double *l1 = dag ? op->uee.data() : op->lee.data();
    const double *l2 = dag ? op->ueem.data() : op->leem.data();
    const double *u1 = dag ? op->lee.data() : op->uee.data();
    const double *u2 = dag ? op->leem.data() : op->ueem.data();
    for (int col = 0; col < Ls; col++) {
      std::vector<double> psi(Ls, 0.0), res_v(Ls, 0.0);
      psi[col] = 1.0;
      double tmp = 0, acc = 0, res;
      if (Ls == 1) {
        res_v[0] = psi[0] / op->dee[0];
      } else {
        res = psi[0];
        if (carry) tmp = res; else acc = l2[0] * res;
        res_v[0] = res;
        for (int s = 1; s < Ls - 1; s++) {
          res = psi[s];
          if (carry) res -= l1[s - 1] * tmp;
          if (carry) tmp = res; else acc += l2[s] * res;
          res_v[s] = res;
        }
        res = psi[Ls - 1];
        if (carry) res -= l1[Ls - 2] * tmp;
        if (!carry) res -= acc;
        res = (1.0 / op->dee[Ls - 1]) * res;
        res_v[Ls - 1] = res;
        const double accb = carry ? res : 0.0;
        double tmpb = carry ? 0.0 : res;
        for (int s = Ls - 2; s >= 0; s--) {
          res = (1.0 / op->dee[s]) * res_v[s] - u1[s] * tmpb - u2[s] * accb;
          tmpb = carry ? 0.0 : res;
          res_v[s] = res;
        }
      }
      double *A = out.m.data() + chi * Ls * Ls;
      for (int s = 0; s < Ls; s++) A[s * Ls + col] = res_v[s];
    }
  }
}

static void build_smats(Op *op) {
  const int Ls = op->g.Ls;
  const double mass = op->mass;
  std::vector<double> lower(Ls), diag(Ls), upper(Ls);
  // Meooe5D: diag bs, upper/lower cs, corners -mass*cs (CayleyFermion5DImplementation.h:165-174)
  for (int s = 0; s < Ls; s++) { diag[s] = op->bs[s]; upper[s] = op->cs[s]; lower[s] = op->cs[s]; }
  upper[Ls - 1] = -mass * upper[Ls - 1];
  lower[0] = -mass * lower[0];
  m5d_matrix(Ls, lower, diag, upper, false, op->smat[Op::S_MEOOE5D]);
  // MeooeDag5D (:249-271)
  for (int s = 0; s < Ls; s++) {
    diag[s] = op->bs[s];
    if (Ls == 1) { upper[s] = -mass * op->cs[0]; lower[s] = -mass * op->cs[0]; continue; }
    if (s == 0) { upper[s] = op->cs[s + 1]; lower[s] = -mass * op->cs[Ls - 1]; }
    else if (s == Ls - 1) { upper[s] = -mass * op->cs[0]; lower[s] = op->cs[s - 1]; }
    else { upper[s] = op->cs[s + 1]; lower[s] = op->cs[s - 1]; }
  }
  m5d_matrix(Ls, lower, diag, upper, true, op->smat[Op::S_MEOOEDAG5D]);
  // Mooee (:190-204)
  for (int s = 0; s < Ls; s++) { diag[s] = op->bee[s]; upper[s] = -op->cee[s]; lower[s] = -op->cee[s]; }
  upper[Ls - 1] = -mass * upper[Ls - 1];
  lower[0] = -mass * lower[0];
  m5d_matrix(Ls, lower, diag, upper, false, op->smat[Op::S_MOOEE]);
  // MooeeDag (:205-233)
  for (int s = 0; s < Ls; s++) {
    diag[s] = op->bee[s];
    if (Ls == 1) { upper[s] = mass * op->cee[0]; lower[s] = mass * op->cee[0]; continue; }
    if (s == 0) { upper[s] = -op->cee[s + 1]; lower[s] = mass * op->cee[Ls - 1]; }
    else if (s == Ls - 1) { upper[s] = mass * op->cee[0]; lower[s] = -op->cee[s - 1]; }
    else { upper[s] = -op->cee[s + 1]; lower[s] = -op->cee[s - 1]; }
  }
  m5d_matrix(Ls, lower, diag, upper, true, op->smat[Op::S_MOOEEDAG]);
  mooee_inv_matrix(op, false, op->smat[Op::S_MOOEEINV]);
  mooee_inv_matrix(op, true, op->smat[Op::S_MOOEEINVDAG]);
  // M5D / M5Ddag off-diagonal parts (diag 1 applies to phi): upper -1, lower -1, corners +mass
  for (int s = 0; s < Ls; s++) { diag[s] = 0; upper[s] = -1; lower[s] = -1; }
  upper[Ls - 1] = mass;
  lower[0] = mass;
  m5d_matrix(Ls, lower, diag, upper, false, op->smat[Op::S_M5D_OFF]);
  m5d_matrix(Ls, lower, diag, upper, true, op->smat[Op::S_M5DDAG_OFF]);
  // products for the fused Schur operator: (A B)[s][t] = sum_u A[s][u] B[u][t] per chirality
  auto product = [Ls](const SMat &A, const SMat &B, SMat &C) {
    C.m.assign(2 * Ls * Ls, 0.0);
    for (int chi = 0; chi < 2; chi++)
      for (int s = 0; s < Ls; s++)
        for (int t = 0; t < Ls; t++) {
          double v = 0;
          for (int u = 0; u < Ls; u++) v += A.m[chi * Ls * Ls + s * Ls + u] * B.m[chi * Ls * Ls + u * Ls + t];
          C.m[chi * Ls * Ls + s * Ls + t] = v;
        }
  };
  product(op->smat[Op::S_MEOOE5D], op->smat[Op::S_MOOEEINV], op->smat[Op::S_MEO5D_INV]);
  product(op->smat[Op::S_MOOEEINVDAG], op->smat[Op::S_MEOOEDAG5D], op->smat[Op::S_INVDAG_MEODAG5D]);
}

static int upload_smats(Op *op) {
  const int Ls = op->g.Ls;
  const size_t per = 2 * (size_t)Ls * Ls;
  std::vector<double> all(per * Op::S_COUNT);
  for (int i = 0; i < Op::S_COUNT; i++) std::memcpy(&all[i * per], op->smat[i].m.data(), per * sizeof(double));
  std::vector<float> allf(all.begin(), all.end());
  for (int p = 0; p < 2; p++)
    if (op->smat_dev[p]) {
      (void)hipFree(op->smat_dev[p]);
      op->smat_dev[p] = nullptr;
    }
  GX_CHECK_HIP(hipMalloc(&op->smat_dev[GX_F64], all.size() * sizeof(double)));
  GX_CHECK_HIP(hipMemcpy(op->smat_dev[GX_F64], all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
  GX_CHECK_HIP(hipMalloc(&op->smat_dev[GX_F32], allf.size() * sizeof(float)));
  GX_CHECK_HIP(hipMemcpy(op->smat_dev[GX_F32], allf.data(), allf.size() * sizeof(float), hipMemcpyHostToDevice));
  return GX_OK;
}

// Per-slice Schur coefficients (see Op::coef_dev): built on the host in FP64, copied to the device in
// FP64 and FP32, in place when the arrays exist (captured graphs keep their pointers valid).
int op_upload_coefs(Op *op) {
  const int Ls = op->g.Ls;
  std::vector<double> c(4 * (size_t)Ls);
  for (int s = 0; s < Ls; s++) {
    if (op->kind == GX_WILSON) {
      const double m = op->slice_mass.empty() ? op->mass : op->slice_mass[s];
      const double d = 4.0 + m;  // WilsonFermionImplementation.h:71, diag_mass = 4 + m
      c[s] = -1.0 / d;
      c[Ls + s] = d;
      c[2 * Ls + s] = 1.0 / d;
      c[3 * Ls + s] = d;
    } else {
      c[s] = -1.0;
      c[Ls + s] = 1.0;
      c[2 * Ls + s] = 1.0;
      c[3 * Ls + s] = 1.0;
    }
  }
  op->coef_host = c;
  op->coef_host_f.assign(c.begin(), c.end());
  if (!op->coef_dev[0]) GX_CHECK_HIP(hipMalloc(&op->coef_dev[0], c.size() * sizeof(double)));
  if (!op->coef_dev[1]) GX_CHECK_HIP(hipMalloc(&op->coef_dev[1], c.size() * sizeof(float)));
  // copies on the operator's own stream: a synchronous copy would touch the null stream, which other
  // host threads' graph captures (serving slots) forbid
  hipStream_t st = op->ctx->stream;
  GX_CHECK_HIP(hipMemcpyAsync(op->coef_dev[0], op->coef_host.data(), c.size() * sizeof(double),
                              hipMemcpyHostToDevice, st));
  GX_CHECK_HIP(hipMemcpyAsync(op->coef_dev[1], op->coef_host_f.data(), c.size() * sizeof(float),
                              hipMemcpyHostToDevice, st));
  GX_CHECK_HIP(hipStreamSynchronize(st));
  return GX_OK;
}

int op_coef(Op *op, int prec, const void **coef) {
  *coef = op->coef_dev[prec == GX_F64 ? 0 : 1];
  if (!*coef) {
    set_error("operator coefficients missing");
    return GX_EINVAL;
  }
  return GX_OK;
}

// matrices in the arithmetic precision of storage prec (FP16 storage computes in float)
int op_smat(Op *op, int prec, int which, const void **mats) {
  const size_t per = 2 * (size_t)op->g.Ls * op->g.Ls;
  if (prec == GX_F64) *mats = (const double *)op->smat_dev[GX_F64] + which * per;
  else *mats = (const float *)op->smat_dev[GX_F32] + which * per;
  return GX_OK;
}

// ---------------- operator building blocks (half fields) ----------------
int apply_dhop_half(Op *op, Field *in, Field *out, int cb_out, int dag, const Gate *gate) {
  void *Ue, *Uo;
  GX_TRY(op_gauge(op, in->prec, &Ue, &Uo));
  void *U = cb_out == 0 ? Ue : Uo;
  out->cb = cb_out;
  return by_prec(in->prec, [&](auto t) {
    return launch_dhop<decltype(t)>(op->g, in->data, out->data, U, cb_out, dag, gate, op->ctx->stream);
  });
}

static int apply_smat(Op *op, Field *in, Field *phi, Field *out, int which, const Gate *gate) {
  const void *m;
  GX_TRY(op_smat(op, in->prec, which, &m));
  out->cb = in->cb;
  return by_prec(in->prec, [&](auto t) {
    return launch_smat<decltype(t)>(op->g, in->data, phi ? phi->data : nullptr, out->data, m, gate, op->ctx->stream);
  });
}

int apply_meooe(Op *op, Field *in, Field *out, int dag, const Gate *gate) {
  const int cb_out = 1 - in->cb;
  if (op->kind == GX_WILSON) return apply_dhop_half(op, in, out, cb_out, dag, gate);
  Field *t = op_tmp(op, in->prec, 3);
  if (!t) return GX_ENOMEM;
  if (!dag) {
    GX_TRY(apply_smat(op, in, nullptr, t, Op::S_MEOOE5D, gate));
    return apply_dhop_half(op, t, out, cb_out, 0, gate);
  }
  GX_TRY(apply_dhop_half(op, in, t, cb_out, 1, gate));
  return apply_smat(op, t, nullptr, out, Op::S_MEOOEDAG5D, gate);
}

// which: 0 Mooee, 1 MooeeDag, 2 MooeeInv, 3 MooeeInvDag
int apply_mooee(Op *op, Field *in, Field *out, int which, const Gate *gate) {
  if (op->kind == GX_WILSON) {
    if (!op->slice_mass.empty()) {  // d_s (Mooee) or 1/d_s (MooeeInv) on each slice
      const void *c;
      GX_TRY(op_coef(op, in->prec, &c));
      const size_t off = (size_t)(which < 2 ? 3 : 2) * op->g.Ls;
      const void *cs = in->prec == GX_F64 ? (const void *)((const double *)c + off) : (const void *)((const float *)c + off);
      out->cb = in->cb;
      return by_prec(in->prec, [&](auto t) {
        return launch_scale_slices<decltype(t)>(op->g, in->data, out->data, cs, gate, op->ctx->stream);
      });
    }
    const double d = 4.0 + op->mass;
    return blas_axpby(out, which < 2 ? d : 1.0 / d, in, 0.0, nullptr, gate);
  }
  static const int map[4] = {Op::S_MOOEE, Op::S_MOOEEDAG, Op::S_MOOEEINV, Op::S_MOOEEINVDAG};
  return apply_smat(op, in, nullptr, out, map[which], gate);
}

// SchurDiagMooeeOperator::Mpc / MpcDag, LinearOperator.h:330-348
// Fused Schur operator (same operator as below, fewer passes over the fields):
//   Wilson  Mpc    = d in - (1/d) Dhop(Dhop(in)),  d = 4 + m                      2 kernels
//   Cayley  Mpc    = Mooee in - Dhop(P1 Dhop(Meooe5D in)),  P1 = Meooe5D MooeeInv   3 kernels
//           MpcDag = MooeeDag in - MeooeDag5D Dhop^dag(P2 Dhop^dag in),
//                    P2 = MooeeInvDag MeooeDag5D                                   3 kernels
// The fifth-dimension matrices act across each site's s-row in the hop kernel's epilogue.
template <typename T>
static int mpc_fused(Op *op, Field *in, Field *out, int dag, const Gate *gate, const HopFuse *fz) {
  const int prec = in->prec, cb_in = in->cb, cb_mid = 1 - cb_in;
  void *Ue, *Uo;
  GX_TRY(op_gauge(op, prec, &Ue, &Uo));
  void *U_mid = cb_mid == 0 ? Ue : Uo, *U_in = cb_in == 0 ? Ue : Uo;
  hipStream_t st = op->ctx->stream;
  const Geom &g = op->g;
  Field *t0 = op_tmp(op, prec, 0), *t1 = op_tmp(op, prec, 1), *t3 = op_tmp(op, prec, 3);
  if (!t0 || !t1 || !t3) return GX_ENOMEM;
  const void *coef;  // per-slice (alpha_s, beta_s) of the EPI 1 hops: Wilson (-1/d_s, d_s), Cayley (-1, 1)
  GX_TRY(op_coef(op, prec, &coef));
  if (op->kind == GX_WILSON) {
    GX_TRY(launch_dhop_epi<T>(g, in->data, t0->data, U_mid, cb_mid, dag, gate, st, 0, nullptr, nullptr, 1.0, 0.0));
    GX_TRY(launch_dhop_epi<T>(g, t0->data, out->data, U_in, cb_in, dag, gate, st, 1, in->data, nullptr, 0.0, 0.0, 0,
                              coef, fz));
  } else if (!dag) {
    const void *m_meo, *m_mooee, *m_p1;
    GX_TRY(op_smat(op, prec, Op::S_MEOOE5D, &m_meo));
    GX_TRY(op_smat(op, prec, Op::S_MOOEE, &m_mooee));
    GX_TRY(op_smat(op, prec, Op::S_MEO5D_INV, &m_p1));
    // t0 = Meooe5D in, t1 = Mooee in (one read of in)
    GX_TRY(launch_smat2<T>(g, in->data, nullptr, 0.0, t0->data, m_meo, t1->data, m_mooee, gate, st));
    GX_TRY(launch_dhop_epi<T>(g, t0->data, t3->data, U_mid, cb_mid, 0, gate, st, 2, nullptr, m_p1, 1.0, 0.0));
    GX_TRY(launch_dhop_epi<T>(g, t3->data, out->data, U_in, cb_in, 0, gate, st, 1, t1->data, nullptr, 0.0, 0.0, 0,
                              coef, fz));
  } else {
    const void *m_mooeedag, *m_p2, *m_meodag;
    GX_TRY(op_smat(op, prec, Op::S_MOOEEDAG, &m_mooeedag));
    GX_TRY(op_smat(op, prec, Op::S_INVDAG_MEODAG5D, &m_p2));
    GX_TRY(op_smat(op, prec, Op::S_MEOOEDAG5D, &m_meodag));
    GX_TRY(launch_smat2<T>(g, in->data, nullptr, 0.0, t1->data, m_mooeedag, nullptr, nullptr, gate, st));
    GX_TRY(launch_dhop_epi<T>(g, in->data, t3->data, U_mid, cb_mid, 1, gate, st, 2, nullptr, m_p2, 1.0, 0.0));
    GX_TRY(launch_dhop_epi<T>(g, t3->data, out->data, U_in, cb_in, 1, gate, st, 2, t1->data, m_meodag, -1.0, 1.0, 0,
                              nullptr, fz));
  }
  out->cb = cb_in;
  return GX_OK;
}

int apply_mpc(Op *op, Field *in, Field *out, int dag, const Gate *gate, const HopFuse *fz) {
  if (fz && fz->nparts) *fz->nparts = 0;
  if (tune_get(GX_TUNE_FUSED, in->prec))
    return by_prec(in->prec, [&](auto t) { return mpc_fused<decltype(t)>(op, in, out, dag, gate, fz); });
  Field *t1 = op_tmp(op, in->prec, 0), *t2 = op_tmp(op, in->prec, 1);
  if (!t1 || !t2) return GX_ENOMEM;
  GX_TRY(apply_meooe(op, in, t1, dag, gate));
  GX_TRY(apply_mooee(op, t1, t2, dag ? 3 : 2, gate));
  GX_TRY(apply_meooe(op, t2, t1, dag, gate));
  GX_TRY(apply_mooee(op, in, out, dag ? 1 : 0, gate));
  GX_TRY(blas_axpby(out, 1.0, out, -1.0, t1, gate));
  out->cb = in->cb;
  return GX_OK;
}

int apply_mpcdagmpc(Op *op, Field *in, Field *out, const Gate *gate) {
  Field *t = op_tmp(op, in->prec, 2);
  if (!t) return GX_ENOMEM;
  GX_TRY(apply_mpc(op, in, t, 0, gate));
  return apply_mpc(op, t, out, 1, gate);
}

// host (lexicographic, s innermost, AoS) <-> device (EO split, 64-entry blocks) index maps; the
// stage holds 12 columns per half (half_pad entries each, padding zero) in the grouping of the
// storage precision with the stage's type (FP64: 1 component per 16 bytes, FP32: 2); the FP16
// packing reads an FP32-grouped stage
template <typename T>
static int vw_of() {
  return 16 / (int)(2 * sizeof(T));
}
template <typename T>
static void gather_to_device_layout(const Field *f, const double *host, T *stage) {
  const int vw = vw_of<T>();
  const Geom &g = f->g;
  const int Ls = g.Ls;
  const int64_t hp = 12 * g.half_pad();
  if (f->subset == GX_FULL) {
    parallel_for(g.V4, [&](int64_t b, int64_t e) {
      for (int64_t s4 = b; s4 < e; s4++) {
        int x[4];
        coor_lex(g, s4, x);
        const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
        const int64_t h = half_index(g, x);
        for (int s = 0; s < Ls; s++)
          for (int k = 0; k < 12; k++) {
            const double *src = host + ((s4 * Ls + s) * 12 + k) * 2;
            T *dst = stage + (cb * hp + blk_index(h * Ls + s, k, 12, vw)) * 2;
            dst[0] = (T)src[0];
            dst[1] = (T)src[1];
          }
      }
    });
  } else {
    parallel_for(g.H4, [&](int64_t b, int64_t e) {
      for (int64_t h = b; h < e; h++)
        for (int s = 0; s < Ls; s++)
          for (int k = 0; k < 12; k++) {
            const double *src = host + ((h * Ls + s) * 12 + k) * 2;
            T *dst = stage + blk_index(h * Ls + s, k, 12, vw) * 2;
            dst[0] = (T)src[0];
            dst[1] = (T)src[1];
          }
    });
  }
}

template <typename T>
static void scatter_from_device_layout(const Field *f, const T *stage, double *host) {
  const int vw = vw_of<T>();
  const Geom &g = f->g;
  const int Ls = g.Ls;
  const int64_t hp = 12 * g.half_pad();
  if (f->subset == GX_FULL) {
    parallel_for(g.V4, [&](int64_t b, int64_t e) {
      for (int64_t s4 = b; s4 < e; s4++) {
        int x[4];
        coor_lex(g, s4, x);
        const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
        const int64_t h = half_index(g, x);
        for (int s = 0; s < Ls; s++)
          for (int k = 0; k < 12; k++) {
            double *dst = host + ((s4 * Ls + s) * 12 + k) * 2;
            const T *src = stage + (cb * hp + blk_index(h * Ls + s, k, 12, vw)) * 2;
            dst[0] = (double)src[0];
            dst[1] = (double)src[1];
          }
      }
    });
  } else {
    parallel_for(g.H4, [&](int64_t b, int64_t e) {
      for (int64_t h = b; h < e; h++)
        for (int s = 0; s < Ls; s++)
          for (int k = 0; k < 12; k++) {
            double *dst = host + ((h * Ls + s) * 12 + k) * 2;
            const T *src = stage + blk_index(h * Ls + s, k, 12, vw) * 2;
            dst[0] = (double)src[0];
            dst[1] = (double)src[1];
          }
    });
  }
}

int apply_full_m(Op *op, Field *in, Field *out, int dag) {
  for (int cb = 0; cb < 2; cb++) {
    Field in_other = half_view(in, 1 - cb), in_same = half_view(in, cb), o = half_view(out, cb);
    GX_TRY(apply_meooe(op, &in_other, &o, dag));
    Field *t = op_tmp(op, in->prec, 1);
    if (!t) return GX_ENOMEM;
    t->cb = cb;
    GX_TRY(apply_mooee(op, &in_same, t, dag ? 1 : 0));
    GX_TRY(blas_axpby(&o, 1.0, &o, 1.0, t));
  }
  return GX_OK;
}

int slice_transfer(Field *lo, Field *hi, int s, bool extract) {
  bool ok = lo && hi && (lo->subset == GX_FULL) == (hi->subset == GX_FULL) && lo->g.Ls == 1 && lo->prec == hi->prec &&
            (lo->prec == GX_F64 || lo->prec == GX_F32) && s >= 0 && s < hi->g.Ls;
  for (int d = 0; d < 4 && ok; d++) ok = lo->g.L[d] == hi->g.L[d];
  if (!ok) {
    set_error("slice: needs two full or two half fields of one precision on one 4D lattice, the 4D one with Ls = 1, "
              "and 0 <= s < Ls");
    return GX_EINVAL;
  }
  const int nhalf = lo->subset == GX_FULL ? 2 : 1;
  if (nhalf == 1) (extract ? lo : hi)->cb = (extract ? hi : lo)->cb;  // a half field takes its source's tag
  hipStream_t st = hi->ctx->stream;
  const int Ls = hi->g.Ls;
  if (lo->prec == GX_F64) {
    if (extract) GX_TRY(launch_slice_proj<double>(hi->g, hi->data, Ls, s, lo->data, 1, 0, 0xF, 0, st, nhalf));
    else GX_TRY(launch_slice_proj<double>(hi->g, lo->data, 1, 0, hi->data, Ls, s, 0xF, 0, st, nhalf));
  } else {
    if (extract) GX_TRY(launch_slice_proj<float>(hi->g, hi->data, Ls, s, lo->data, 1, 0, 0xF, 0, st, nhalf));
    else GX_TRY(launch_slice_proj<float>(hi->g, lo->data, 1, 0, hi->data, Ls, s, 0xF, 0, st, nhalf));
  }
  return GX_OK;
}

}  // namespace gx

using namespace gx;


static Field *F(gx_field f) { return static_cast<Field *>(f); }
static Op *O(gx_op o) { return static_cast<Op *>(o); }

#define GX_REQUIRE(cond, code, ...) \
  do {                              \
    if (!(cond)) {                  \
      set_error(__VA_ARGS__);       \
      return code;                  \
    }                               \
  } while (0)

extern "C" {

const char *gx_last_error(void) { return g_err; }
const char *gx_version(void) { return "gx 0.1 (gfx950)"; }

int gx_tune_set(int knob, int value) {
  for (int p = GX_F64; p <= GX_F16; p++) GX_TRY(gx_tune_set_prec(knob, p, value));
  return GX_OK;
}

int gx_tune_set_prec(int knob, int prec, int value) {
  GX_REQUIRE(knob >= 0 && knob < GX_TUNE_COUNT, GX_EINVAL, "unknown tuning knob %d", knob);
  GX_REQUIRE(prec >= GX_F64 && prec <= GX_F16, GX_EINVAL, "unknown precision %d", prec);
  if (knob == GX_TUNE_DHOP_MAP) GX_REQUIRE(value >= 0 && value <= 4, GX_EINVAL, "tile map takes 0..4");
  else if (knob <= GX_TUNE_FUSED) GX_REQUIRE(value == 0 || value == 1, GX_EINVAL, "knob %d takes 0 or 1", knob);
  else if (knob == GX_TUNE_DHOP_SHAPE) GX_REQUIRE(value >= 0 && value <= 8, GX_EINVAL, "tile shape takes 0..8 waves");
  else if (knob == GX_TUNE_DHOP_PERSIST || knob == GX_TUNE_DHOP_MFMA || knob == GX_TUNE_DHOP_TILE)
    GX_REQUIRE(value == 0 || value == 1, GX_EINVAL, "knob %d takes 0 or 1", knob);
  g_tune[prec][knob] = value;
  g_tune_version++;
  return GX_OK;
}

int gx_tune_get(int knob, int *value) { return gx_tune_get_prec(knob, GX_F64, value); }

int gx_tune_get_prec(int knob, int prec, int *value) {
  GX_REQUIRE(knob >= 0 && knob < GX_TUNE_COUNT && value, GX_EINVAL, "unknown tuning knob %d", knob);
  GX_REQUIRE(prec >= GX_F64 && prec <= GX_F16, GX_EINVAL, "unknown precision %d", prec);
  *value = g_tune[prec][knob];
  return GX_OK;
}

int gx_ctx_create(int device, gx_ctx *out) {
  GX_REQUIRE(out, GX_EINVAL, "null output");
  GX_CHECK_HIP(hipSetDevice(device));
  gx_ctx c = new gx_ctx_s();
  c->device = device;
  GX_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  for (int i = 0; i < 16; i++) GX_CHECK_HIP(hipEventCreate(&c->ev[i]));
  GX_CHECK_HIP(hipMalloc(&c->red_partial, sizeof(double) * kRedBlocks * kMaxRhs));
  GX_CHECK_HIP(hipMalloc(&c->digest, sizeof(unsigned long long)));
  GX_CHECK_HIP(hipMalloc(&c->red_result, sizeof(double) * kRedSlots));
  GX_CHECK_HIP(hipHostMalloc(&c->pinned, sizeof(double) * kRedSlots, hipHostMallocDefault));
  *out = c;
  return GX_OK;
}

int gx_ctx_destroy(gx_ctx c) {
  if (!c) return GX_OK;
  (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 16; i++) (void)hipEventDestroy(c->ev[i]);
  (void)hipFree(c->red_partial);
  (void)hipFree(c->digest);
  (void)hipFree(c->red_result);
  (void)hipHostFree(c->pinned);
  (void)hipStreamDestroy(c->stream);
  delete c;
  return GX_OK;
}

// WilsonLoops<Gimpl>::avgPlaquette / linkTrace (WilsonLoops.h:121-126,203-219) of a host gauge field
int gx_gauge_statistics(gx_ctx c, const int dims[4], const double *U, double *plaquette, double *link_trace) {
  GX_REQUIRE(c && dims && U, GX_EINVAL, "gauge_statistics: null argument");
  for (int d = 0; d < 4; d++) GX_REQUIRE(dims[d] >= 1, GX_EINVAL, "gauge_statistics: bad extent %d", dims[d]);
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const size_t ub = (size_t)V * 4 * 9 * 2 * sizeof(double);
  const int nparts = 1024;
  void *Ud = nullptr;
  double *part = nullptr;
  hipError_t e = hipMalloc(&Ud, ub);
  if (e == hipSuccess) e = hipMalloc((void **)&part, sizeof(double) * 2 * nparts);
  if (e != hipSuccess) {
    if (Ud) (void)hipFree(Ud);
    set_error("gauge_statistics: %s", hipGetErrorString(e));
    return GX_ENOMEM;
  }
  std::vector<double> host(2 * nparts);
  int rc = GX_OK;
  e = hipMemcpyAsync(Ud, U, ub, hipMemcpyHostToDevice, c->stream);
  if (e != hipSuccess) {
    set_error("gauge_statistics: %s", hipGetErrorString(e));
    rc = GX_EHIP;
  }
  if (rc == GX_OK) rc = launch_gauge_stats(c, dims, Ud, part, nparts, host.data());
  (void)hipFree(Ud);
  (void)hipFree(part);
  if (rc != GX_OK) return rc;
  if (plaquette) *plaquette = host[0] / (double)V / 6.0 / 3.0;
  if (link_trace) *link_trace = host[1] / (double)V / 12.0;
  return GX_OK;
}

int gx_ctx_sync(gx_ctx c) {
  GX_REQUIRE(c, GX_EINVAL, "null ctx");
  GX_CHECK_HIP(hipStreamSynchronize(c->stream));
  return GX_OK;
}

int gx_ctx_event_record(gx_ctx c, int slot) {
  GX_REQUIRE(c && slot >= 0 && slot < 16, GX_EINVAL, "bad event slot");
  GX_CHECK_HIP(hipEventRecord(c->ev[slot], c->stream));
  return GX_OK;
}

int gx_ctx_event_elapsed(gx_ctx c, int s0, int s1, double *ms) {
  GX_REQUIRE(c && s0 >= 0 && s0 < 16 && s1 >= 0 && s1 < 16 && ms, GX_EINVAL, "bad event slot");
  GX_CHECK_HIP(hipEventSynchronize(c->ev[s1]));
  float f = 0;
  GX_CHECK_HIP(hipEventElapsedTime(&f, c->ev[s0], c->ev[s1]));
  *ms = f;
  return GX_OK;
}

// ---------------- fields ----------------
int gx_field_create(gx_ctx ctx, const int dims[4], int Ls, int prec, int subset, gx_field *out) {
  GX_REQUIRE(ctx && out && geom_ok(dims, Ls), GX_EINVAL, "bad field geometry");
  GX_REQUIRE(subset >= -1 && subset <= 1, GX_EINVAL, "bad subset %d", subset);
  Field *f;
  GX_TRY(field_create(ctx, make_geom(dims, Ls), prec, subset, &f));
  gx_field g = new gx_field_s();
  *static_cast<Field *>(g) = *f;
  delete f;
  *out = g;
  return GX_OK;
}

int gx_field_destroy(gx_field f) {
  if (!f) return GX_OK;
  if (f->data) (void)hipFree(f->data);
  delete f;
  return GX_OK;
}

// FP16 storage <-> FP32 device layout on the host: per half block, 12 columns of binary16 pairs
// scaled by each site's largest |component| (column 12, float).
static void pack_f16(const Field *f, const float *in, uint32_t *out) {
  const int64_t P = f->g.half_pad(), nh = f->halves();
  for (int64_t hf = 0; hf < nh; hf++)
    parallel_for(P, [&](int64_t b, int64_t e) {
      const float *ih = in + hf * 12 * P * 2;
      uint32_t *oh = out + hf * 13 * P;
      for (int64_t i = b; i < e; i++) {
        float m = 0.f;
        for (int k = 0; k < 12; k++) {
          const float *v = ih + blk_index(i, k, 12, 2) * 2;
          m = std::max(m, std::max(std::fabs(v[0]), std::fabs(v[1])));
        }
        // a subnormal maximum is scaled by 2^64 (exact) first, as 1/m would overflow
        const float up = (m > 0.f && m < FLT_MIN) ? 0x1p64f : 1.f;
        const float inv = m > 0.f ? 1.f / (m * up) : 0.f;
        for (int k = 0; k < 12; k++) {
          const float *v = ih + blk_index(i, k, 12, 2) * 2;
          const half_t pr[2] = {(half_t)((v[0] * up) * inv), (half_t)((v[1] * up) * inv)};
          std::memcpy(&oh[blk_index(i, k, 13, 4)], pr, 4);
        }
        std::memcpy(&oh[blk_index(i, 12, 13, 4)], &m, 4);
      }
    });
}

static void unpack_f16(const Field *f, const uint32_t *in, float *out) {
  const int64_t P = f->g.half_pad(), nh = f->halves();
  for (int64_t hf = 0; hf < nh; hf++)
    parallel_for(P, [&](int64_t b, int64_t e) {
      const uint32_t *ih = in + hf * 13 * P;
      float *oh = out + hf * 12 * P * 2;
      for (int64_t i = b; i < e; i++) {
        float m;
        std::memcpy(&m, &ih[blk_index(i, 12, 13, 4)], 4);
        for (int k = 0; k < 12; k++) {
          half_t pr[2];
          std::memcpy(pr, &ih[blk_index(i, k, 13, 4)], 4);
          float *v = oh + blk_index(i, k, 12, 2) * 2;
          v[0] = (float)pr[0] * m;
          v[1] = (float)pr[1] * m;
        }
      }
    });
}

int gx_field_import(gx_field f, const double *host) {
  GX_REQUIRE(f && host, GX_EINVAL, "null argument");
  if (f->prec == GX_F16) {
    std::vector<float> stage((size_t)f->stored() * 2, 0.f);
    gather_to_device_layout<float>(f, host, stage.data());
    std::vector<uint32_t> packed(f->bytes / 4);
    pack_f16(f, stage.data(), packed.data());
    GX_CHECK_HIP(hipMemcpyAsync(f->data, packed.data(), f->bytes, hipMemcpyHostToDevice, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
  } else if (f->prec == GX_F64) {
    std::vector<double> stage((size_t)f->stored() * 2, 0.0);
    gather_to_device_layout<double>(f, host, stage.data());
    GX_CHECK_HIP(hipMemcpyAsync(f->data, stage.data(), f->bytes, hipMemcpyHostToDevice, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
  } else {
    std::vector<float> stage((size_t)f->stored() * 2, 0.f);
    gather_to_device_layout<float>(f, host, stage.data());
    GX_CHECK_HIP(hipMemcpyAsync(f->data, stage.data(), f->bytes, hipMemcpyHostToDevice, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
  }
  return GX_OK;
}

int gx_field_gaussian_slices(gx_field f, uint64_t seed, int s0, int ns) {
  GX_REQUIRE(f && f->subset == GX_FULL && (f->prec == GX_F64 || f->prec == GX_F32), GX_EINVAL,
             "field_gaussian: needs a full F64 or F32 field");
  GX_REQUIRE(s0 >= 0 && ns >= 1 && s0 + ns <= f->g.Ls, GX_EINVAL, "field_gaussian: slices %d..%d outside Ls = %d",
             s0, s0 + ns - 1, f->g.Ls);
  // the host generator's key schedule (SplitMix64 of an affine map of the seed)
  uint64_t x = seed * 0x9E6C63D0676A9A99ULL + 0xBB67AE8584CAA73BULL + 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  const uint64_t key = x ^ (x >> 31);
  if (f->prec == GX_F64) GX_TRY(launch_gaussian<double>(f->g, f->data, key, f->ctx->stream, s0, ns));
  else GX_TRY(launch_gaussian<float>(f->g, f->data, key, f->ctx->stream, s0, ns));
  f->cb = 0;
  return GX_OK;
}

int gx_field_gaussian(gx_field f, uint64_t seed) {
  GX_REQUIRE(f, GX_EINVAL, "null field");
  return gx_field_gaussian_slices(f, seed, 0, f->g.Ls);
}

int gx_field_point_source(gx_field f, const int x[4], int spin, int colour) {
  GX_REQUIRE(f && x && f->subset == GX_FULL && (f->prec == GX_F64 || f->prec == GX_F32), GX_EINVAL,
             "point_source: needs a full F64 or F32 field");
  const bool batch = spin == -1 && colour == -1;
  GX_REQUIRE(batch ? f->g.Ls <= 12 : (spin >= 0 && spin < 4 && colour >= 0 && colour < 3), GX_EINVAL,
             "point_source: spin/colour out of range (or Ls > 12 for the spin-colour batch)");
  for (int d = 0; d < 4; d++) GX_REQUIRE(x[d] >= 0 && x[d] < f->g.L[d], GX_EINVAL, "point_source: site outside");
  hipStream_t st = f->ctx->stream;
  GX_CHECK_HIP(hipMemsetAsync(f->data, 0, f->bytes, st));
  const int cb = (x[0] + x[1] + x[2] + x[3]) & 1;
  const int64_t h = half_index(f->g, x);
  const double one64[2] = {1.0, 0.0};
  const float one32[2] = {1.0f, 0.0f};
  const void *one = f->prec == GX_F64 ? (const void *)one64 : (const void *)one32;
  for (int s = 0; s < f->g.Ls; s++) {
    const int k = batch ? s : spin * 3 + colour;
    const int64_t e = (int64_t)cb * 12 * f->g.half_pad() + blk_index(h * f->g.Ls + s, k, 12, f->prec == GX_F64 ? 1 : 2);
    GX_CHECK_HIP(hipMemcpyAsync((char *)f->data + e * f->csize(), one, f->csize(), hipMemcpyHostToDevice, st));
  }
  GX_CHECK_HIP(hipStreamSynchronize(st));
  f->cb = 0;
  return GX_OK;
}

int gx_extract_slice(gx_field out4d, gx_field in5d, int s) { return slice_transfer(out4d, in5d, s, true); }
int gx_insert_slice(gx_field in4d, gx_field out5d, int s) { return slice_transfer(in4d, out5d, s, false); }

int gx_field_digest(gx_field f, uint64_t *digest) {
  GX_REQUIRE(f && digest, GX_EINVAL, "null argument");
  GX_REQUIRE(f->bytes % 8 == 0, GX_EINVAL, "field_digest: size not a multiple of 8 bytes");
  return launch_digest(f->ctx, f->data, (int64_t)(f->bytes / 8), digest);
}

int gx_field_digest_slices(gx_field f, int s0, int ns, uint64_t *digest) {
  GX_REQUIRE(f && digest && f->subset == GX_FULL && f->prec == GX_F64, GX_EINVAL,
             "field_digest_slices: needs a full F64 field");
  return launch_digest_slices(f->ctx, f->g, f->data, s0, ns, digest);
}

int gx_field_export(gx_field f, double *host) {
  GX_REQUIRE(f && host, GX_EINVAL, "null argument");
  if (f->prec == GX_F16) {
    std::vector<uint32_t> packed(f->bytes / 4);
    GX_CHECK_HIP(hipMemcpyAsync(packed.data(), f->data, f->bytes, hipMemcpyDeviceToHost, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
    std::vector<float> stage((size_t)f->stored() * 2);
    unpack_f16(f, packed.data(), stage.data());
    scatter_from_device_layout<float>(f, stage.data(), host);
  } else if (f->prec == GX_F64) {
    std::vector<double> stage((size_t)f->stored() * 2);
    GX_CHECK_HIP(hipMemcpyAsync(stage.data(), f->data, f->bytes, hipMemcpyDeviceToHost, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
    scatter_from_device_layout<double>(f, stage.data(), host);
  } else {
    std::vector<float> stage((size_t)f->stored() * 2);
    GX_CHECK_HIP(hipMemcpyAsync(stage.data(), f->data, f->bytes, hipMemcpyDeviceToHost, f->ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(f->ctx->stream));
    scatter_from_device_layout<float>(f, stage.data(), host);
  }
  return GX_OK;
}

int gx_field_checkerboard(gx_field f, int *cb) {
  GX_REQUIRE(f && cb, GX_EINVAL, "null argument");
  *cb = f->subset == GX_FULL ? GX_FULL : f->cb;
  return GX_OK;
}

int gx_field_set_checkerboard_tag(gx_field f, int cb) {
  GX_REQUIRE(f && f->subset != GX_FULL && (cb == 0 || cb == 1), GX_EINVAL, "tag applies to half fields");
  f->cb = cb;
  return GX_OK;
}

int gx_field_zero(gx_field f) {
  GX_REQUIRE(f, GX_EINVAL, "null field");
  GX_CHECK_HIP(hipMemsetAsync(f->data, 0, f->bytes, f->ctx->stream));
  return GX_OK;
}

int gx_field_copy(gx_field dst, gx_field src) {
  GX_REQUIRE(conformable(F(dst), F(src)), GX_EINVAL, "nonconformable copy");
  GX_CHECK_HIP(hipMemcpyAsync(dst->data, src->data, src->bytes, hipMemcpyDeviceToDevice, src->ctx->stream));
  dst->cb = src->cb;
  return GX_OK;
}

int gx_pick_checkerboard(int cb, gx_field half, gx_field full) {
  GX_REQUIRE(half && full && half->subset != GX_FULL && full->subset == GX_FULL && same_geom(half->g, full->g) &&
                 half->prec == full->prec && (cb == 0 || cb == 1),
             GX_EINVAL, "pickCheckerboard: nonconformable");
  GX_CHECK_HIP(hipMemcpyAsync(half->data, full->half(cb), half->bytes, hipMemcpyDeviceToDevice, full->ctx->stream));
  half->cb = cb;
  return GX_OK;
}

int gx_set_checkerboard(gx_field full, gx_field half) {
  GX_REQUIRE(half && full && half->subset != GX_FULL && full->subset == GX_FULL && same_geom(half->g, full->g) &&
                 half->prec == full->prec,
             GX_EINVAL, "setCheckerboard: nonconformable");
  GX_CHECK_HIP(hipMemcpyAsync(full->half(half->cb), half->data, half->bytes, hipMemcpyDeviceToDevice, full->ctx->stream));
  return GX_OK;
}

int gx_precision_change(gx_field out, gx_field in) {
  GX_REQUIRE(out && in && same_geom(out->g, in->g) && (out->subset == GX_FULL) == (in->subset == GX_FULL), GX_EINVAL,
             "precisionChange: nonconformable");
  return convert_field(F(out), F(in));
}

// ---------------- BLAS1 ----------------
int gx_axpy(gx_field z, double a, gx_field x, gx_field y) {
  GX_REQUIRE(conformable(F(z), F(x)) && conformable(F(x), F(y)), GX_EINVAL, "axpy: nonconformable");
  return blas_axpby(F(z), a, F(x), 1.0, F(y));
}
int gx_axpby(gx_field z, double a, double b, gx_field x, gx_field y) {
  GX_REQUIRE(conformable(F(z), F(x)) && conformable(F(x), F(y)), GX_EINVAL, "axpby: nonconformable");
  return blas_axpby(F(z), a, F(x), b, F(y));
}
int gx_scale(gx_field z, double a, gx_field x) {
  GX_REQUIRE(conformable(F(z), F(x)), GX_EINVAL, "scale: nonconformable");
  return blas_axpby(F(z), a, F(x), 0.0, nullptr);
}
int gx_axpy_norm(gx_field z, double a, gx_field x, gx_field y, double *nrm) {
  GX_REQUIRE(conformable(F(z), F(x)) && conformable(F(x), F(y)) && nrm, GX_EINVAL, "axpy_norm: nonconformable");
  return blas_axpy_norm(F(z), a, F(x), F(y), nrm);
}
int gx_norm2(gx_field x, double *nrm) {
  GX_REQUIRE(x && nrm, GX_EINVAL, "null argument");
  return blas_norm2(F(x), nrm);
}
int gx_inner_product(gx_field x, gx_field y, double *re, double *im) {
  GX_REQUIRE(conformable(F(x), F(y)) && re, GX_EINVAL, "innerProduct: nonconformable");
  return blas_inner(F(x), F(y), re, im);
}
int gx_inner_product_norm(gx_field x, gx_field y, double *re, double *im, double *nrm) {
  GX_REQUIRE(conformable(F(x), F(y)) && re && nrm, GX_EINVAL, "innerProductNorm: nonconformable");
  GX_TRY(blas_inner(F(x), F(y), re, im));
  return blas_norm2(F(x), nrm);
}

// ---------------- operators ----------------
static int op_common(gx_ctx ctx, const int dims[4], int Ls, double mass, const double *U, const double *phases,
                     Op *op) {
  op->ctx = ctx;
  op->g = make_geom(dims, Ls);
  op->mass = mass;
  if (phases)
    for (int i = 0; i < 8; i++) op->phases[i] = phases[i];
  op->host_U.assign(U, U + (size_t)op->g.V4 * 4 * 18);
  op->epoch++;
  GX_TRY(op_upload_coefs(op));
  return upload_gauge(op);
}

int gx_wilson_create(gx_ctx ctx, const int dims[4], int Ls, double mass, const double *U, const double *phases,
                     gx_op *out) {
  GX_REQUIRE(ctx && out && U && geom_ok(dims, Ls), GX_EINVAL, "wilson: bad arguments");
  gx_op op = new gx_op_s();
  op->kind = GX_WILSON;
  int rc = op_common(ctx, dims, Ls, mass, U, phases, op);
  if (rc != GX_OK) {
    gx_op_destroy(op);
    return rc;
  }
  *out = op;
  return GX_OK;
}

int gx_mobius_create(gx_ctx ctx, const int dims[4], int Ls, double mass, double M5, double b, double c,
                     const double *U, const double *phases, gx_op *out) {
  GX_REQUIRE(ctx && out && U && geom_ok(dims, Ls) && Ls >= 2, GX_EINVAL, "mobius: bad arguments");
  gx_op op = new gx_op_s();
  op->kind = GX_MOBIUS;
  op->M5 = M5;
  op->b = b;
  op->c = c;
  int rc = op_common(ctx, dims, Ls, mass, U, phases, op);
  if (rc == GX_OK) {
    cayley_coeffs(op);
    build_smats(op);
    rc = upload_smats(op);
  }
  if (rc != GX_OK) {
    gx_op_destroy(op);
    return rc;
  }
  *out = op;
  return GX_OK;
}

int gx_op_destroy(gx_op op) {
  if (!op) return GX_OK;
  {
    // a base whose links clones still read stays alive (and its links with it)
    std::lock_guard<std::mutex> lk(g_share_mu);
    GX_REQUIRE(op->clones == 0, GX_EINVAL, "op_destroy: %d clone(s) still read this operator's links", op->clones);
    if (op->link_owner) op->link_owner->clones--;
    op->link_owner = nullptr;
  }
  (void)hipStreamSynchronize(op->ctx->stream);
  solver_work_free(op->work);
  free_gauge(op);
  for (int p = 0; p < 3; p++) {
    if (op->smat_dev[p]) (void)hipFree(op->smat_dev[p]);
    for (int i = 0; i < Op::kTmp; i++) field_destroy(op->tmp[p][i]);
  }
  for (int p = 0; p < 2; p++)
    if (op->coef_dev[p]) (void)hipFree(op->coef_dev[p]);
  field_destroy(op->full_tmp);
  delete op;
  return GX_OK;
}

// Links of a base Wilson operator in every precision a solve may use, ready on base's stream.
static int share_links(Op *op, Op *base) {
  std::lock_guard<std::mutex> lk(g_share_mu);  // base is not a clone: op_gauge takes no lock here
  void *ue, *uo;
  GX_TRY(op_gauge(base, GX_F32, &ue, &uo));
  GX_TRY(op_gauge(base, GX_F16, &ue, &uo));
  GX_CHECK_HIP(hipStreamSynchronize(base->ctx->stream));
  for (int p = 0; p < 3; p++)
    for (int cb = 0; cb < 2; cb++) op->U[p][cb] = base->U[p][cb];
  if (op->link_owner) op->link_owner->clones--;
  base->clones++;
  op->link_owner = base;
  op->owns_links = false;
  for (int i = 0; i < 8; i++) op->phases[i] = base->phases[i];
  op->epoch++;
  return GX_OK;
}

int gx_wilson_clone(gx_op base, gx_ctx ctx, int Ls, double mass, gx_op *out) {
  GX_REQUIRE(base && ctx && out && base->kind == GX_WILSON, GX_EINVAL, "wilson_clone: needs a Wilson base operator");
  GX_REQUIRE(!base->link_owner, GX_EINVAL, "wilson_clone: base is itself a clone");
  GX_REQUIRE(Ls >= 1 && Ls <= kMaxRhs, GX_EINVAL, "wilson_clone: Ls=%d out of range (1..%d)", Ls, kMaxRhs);
  gx_op op = new gx_op_s();
  op->kind = GX_WILSON;
  op->ctx = ctx;
  op->g = make_geom(base->g.L, Ls);
  op->mass = mass;
  int rc = op_upload_coefs(op);
  if (rc == GX_OK) rc = share_links(op, base);
  if (rc != GX_OK) {
    for (int p = 0; p < 2; p++)
      if (op->coef_dev[p]) (void)hipFree(op->coef_dev[p]);
    delete op;
    return rc;
  }
  *out = op;
  return GX_OK;
}

int gx_op_share_links(gx_op op, gx_op base) {
  GX_REQUIRE(op && base && op->link_owner && !base->link_owner && base->kind == GX_WILSON, GX_EINVAL,
             "share_links: needs a clone and a Wilson base operator");
  for (int mu = 0; mu < 4; mu++)
    GX_REQUIRE(op->g.L[mu] == base->g.L[mu], GX_EINVAL, "share_links: lattice differs from the base's");
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  return share_links(op, base);
}

int gx_op_footprint(gx_op op, size_t *bytes) {
  GX_REQUIRE(op && bytes, GX_EINVAL, "null argument");
  size_t b = 0;
  const size_t nlink = (size_t)8 * 9 * op->g.H4 * 2;  // complex numbers, both checkerboards
  if (!op->link_owner) {
    if (op->U[GX_F64][0]) b += nlink * 16;
    if (op->U[GX_F32][0]) b += nlink * 8;
    if (op->U[GX_F16][0]) b += nlink * 4;
  }
  for (int p = 0; p < 2; p++)
    if (op->coef_dev[p]) b += (size_t)4 * op->g.Ls * (p == 0 ? 8 : 4);
  for (int p = 0; p < 3; p++) {
    if (op->smat_dev[p]) b += (size_t)Op::S_COUNT * 2 * op->g.Ls * op->g.Ls * (p == GX_F64 ? 8 : 4);
    for (int i = 0; i < Op::kTmp; i++)
      if (op->tmp[p][i]) b += op->tmp[p][i]->bytes;
  }
  if (op->full_tmp) b += op->full_tmp->bytes;
  b += solver_work_bytes(op->work);
  *bytes = b;
  return GX_OK;
}

int gx_op_same_action(gx_op a, gx_op b, int *same) {
  GX_REQUIRE(a && b && same, GX_EINVAL, "null argument");
  // a clone's gauge field is its link owner's
  const Op *ga = a->link_owner ? a->link_owner : O(a), *gb = b->link_owner ? b->link_owner : O(b);
  bool eq = a->kind == b->kind && same_geom(a->g, b->g) && a->mass == b->mass && a->M5 == b->M5 && a->b == b->b &&
            a->c == b->c && a->slice_mass == b->slice_mass && std::memcmp(a->phases, b->phases, sizeof(a->phases)) == 0;
  eq = eq && (ga == gb || (ga->host_U.size() == gb->host_U.size() &&
                           std::memcmp(ga->host_U.data(), gb->host_U.data(), ga->host_U.size() * sizeof(double)) == 0 &&
                           std::memcmp(ga->phases, gb->phases, sizeof(ga->phases)) == 0));
  *same = eq ? 1 : 0;
  return GX_OK;
}

int gx_op_import_gauge(gx_op op, const double *U) {
  GX_REQUIRE(op && U, GX_EINVAL, "null argument");
  GX_REQUIRE(!op->link_owner, GX_EINVAL, "import_gauge: the operator shares another operator's links");
  {
    std::lock_guard<std::mutex> lk(g_share_mu);
    GX_REQUIRE(op->clones == 0, GX_EINVAL, "import_gauge: %d clone(s) still read this operator's links", op->clones);
  }
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  op->host_U.assign(U, U + (size_t)op->g.V4 * 4 * 18);
  op->epoch++;
  return upload_gauge(op);
}

int gx_op_set_mass(gx_op op, double mass) {
  GX_REQUIRE(op, GX_EINVAL, "null op");
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  op->mass = mass;
  op->slice_mass.clear();
  op->epoch++;
  GX_TRY(op_upload_coefs(op));
  if (op->kind == GX_MOBIUS) {
    cayley_coeffs(op);
    build_smats(op);
    return upload_smats(op);
  }
  return GX_OK;
}

int gx_op_set_slice_masses(gx_op op, int n, const double *masses) {
  GX_REQUIRE(op && op->kind == GX_WILSON, GX_EINVAL, "set_slice_masses: needs a Wilson operator");
  GX_REQUIRE(!masses || n == op->g.Ls, GX_EINVAL, "set_slice_masses: %d masses for Ls = %d", n, op->g.Ls);
  if (masses)
    for (int s = 0; s < n; s++) GX_REQUIRE(std::isfinite(masses[s]) && 4.0 + masses[s] != 0.0, GX_EINVAL,
                                           "set_slice_masses: mass %d is %g", s, masses[s]);
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));
  if (masses) op->slice_mass.assign(masses, masses + n);
  else op->slice_mass.clear();
  op->epoch++;
  return op_upload_coefs(op);
}

int gx_op_slice_masses(gx_op op, int n, double *masses) {
  GX_REQUIRE(op && masses && n == op->g.Ls, GX_EINVAL, "slice_masses: needs Ls = %d values", op ? op->g.Ls : 0);
  for (int s = 0; s < n; s++) masses[s] = op->slice_mass.empty() ? op->mass : op->slice_mass[s];
  return GX_OK;
}

static int check_io(Op *op, Field *in, Field *out, bool full) {
  GX_REQUIRE(op && in && out, GX_EINVAL, "null argument");
  GX_REQUIRE(same_geom(in->g, op->g) && same_geom(out->g, op->g), GX_EINVAL, "field geometry does not match operator");
  GX_REQUIRE(in->prec == out->prec, GX_EINVAL, "in/out precision differ");
  GX_REQUIRE(in->prec == GX_F64 || in->prec == GX_F32 || in->prec == GX_F16, GX_EINVAL, "operator precision unsupported");
  GX_REQUIRE((in->subset == GX_FULL) == full && (out->subset == GX_FULL) == full, GX_EINVAL,
             full ? "operator needs full fields" : "operator needs half fields");
  GX_REQUIRE(in != out, GX_EINVAL, "in and out must differ");
  return GX_OK;
}

int gx_M(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), true));
  return apply_full_m(O(op), F(in), F(out), 0);
}
int gx_Mdag(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), true));
  return apply_full_m(O(op), F(in), F(out), 1);
}
int gx_Meooe(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_meooe(O(op), F(in), F(out), 0);
}
int gx_MeooeDag(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_meooe(O(op), F(in), F(out), 1);
}
int gx_Mooee(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mooee(O(op), F(in), F(out), 0);
}
int gx_MooeeDag(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mooee(O(op), F(in), F(out), 1);
}
int gx_MooeeInv(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mooee(O(op), F(in), F(out), 2);
}
int gx_MooeeInvDag(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mooee(O(op), F(in), F(out), 3);
}
int gx_Dhop(gx_op op, gx_field in, gx_field out, int dag) {
  GX_TRY(check_io(O(op), F(in), F(out), true));
  for (int cb = 0; cb < 2; cb++) {
    Field i = half_view(F(in), 1 - cb), o = half_view(F(out), cb);
    GX_TRY(apply_dhop_half(O(op), &i, &o, cb, dag));
  }
  return GX_OK;
}
int gx_DhopEO(gx_op op, gx_field in, gx_field out, int dag) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  GX_REQUIRE(in->cb == GX_ODD, GX_ECB, "DhopEO: input must be on the odd checkerboard");
  return apply_dhop_half(O(op), F(in), F(out), GX_EVEN, dag);
}
int gx_DhopOE(gx_op op, gx_field in, gx_field out, int dag) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  GX_REQUIRE(in->cb == GX_EVEN, GX_ECB, "DhopOE: input must be on the even checkerboard");
  return apply_dhop_half(O(op), F(in), F(out), GX_ODD, dag);
}

// Grid: WilsonFermion(5D)::DhopDir / Mdir (WilsonFermionImplementation.h:345-366): the leg of
// direction dir, displacement disp = +1 (forward neighbour, (1 - gamma_dir), Uds[dir]) or -1 (backward
// neighbour, (1 + gamma_dir), Uds[dir + 4]) of the non-dagger hop, on full F64 fields.
int gx_DhopDir(gx_op op, gx_field in, gx_field out, int dir, int disp) {
  GX_TRY(check_io(O(op), F(in), F(out), true));
  GX_REQUIRE(in->prec == GX_F64, GX_EINVAL, "DhopDir: F64 fields");
  GX_REQUIRE(dir >= 0 && dir < 4 && (disp == 1 || disp == -1), GX_EINVAL, "DhopDir: dir 0..3, disp +-1");
  void *Ue, *Uo;
  GX_TRY(op_gauge(O(op), GX_F64, &Ue, &Uo));
  const bool fwd = disp == 1;
  for (int cb = 0; cb < 2; cb++) {
    Field i = half_view(F(in), 1 - cb), o = half_view(F(out), cb);
    GX_TRY(launch_hop_leg(op->g, dir, fwd, !fwd, 0, cb, i.data, cb == 0 ? Ue : Uo, nullptr, o.data, nullptr,
                          op->ctx->stream));
  }
  return GX_OK;
}

// Grid: DhopDeriv / DhopDerivEO / DhopDerivOE (WilsonFermion5DImplementation.h:211-305,
// WilsonFermionImplementation.h DerivInternal): for every mu, Btilde = Uds[mu](x) P B(x + mu) with
// P = (1 - gamma_mu) (dag = 0) or (1 + gamma_mu) (dag = 1), and mat_mu(x) = sum_{s, spin} Btilde A^dag
// (InsertForce5D).  cbs: the output checkerboards (both for DhopDeriv).  mat: host gauge layout,
// sites of other checkerboards left untouched.
static int dhop_deriv(Op *op, double *mat, Field *A, Field *B, int dag, int cb_first, int cb_last) {
  void *Ue, *Uo;
  GX_TRY(op_gauge(op, GX_F64, &Ue, &Uo));
  const Geom &g = op->g;
  const size_t per = (size_t)g.H4 * 9 * 2 * sizeof(double);
  double *dev = nullptr;
  GX_CHECK_HIP(hipMalloc((void **)&dev, 4 * per));
  std::vector<double> host(4 * per / sizeof(double));
  int rc = GX_OK;
  for (int cb = cb_first; cb <= cb_last && rc == GX_OK; cb++) {
    const Field a = A->subset == GX_FULL ? half_view(A, cb) : *A;
    const Field b = B->subset == GX_FULL ? half_view(B, 1 - cb) : *B;
    for (int mu = 0; mu < 4 && rc == GX_OK; mu++)
      rc = launch_hop_leg(g, mu, 1, dag ? 1 : 0, 1, cb, b.data, cb == 0 ? Ue : Uo, a.data, nullptr,
                          dev + (size_t)mu * g.H4 * 18, op->ctx->stream);
    hipError_t e = rc == GX_OK ? hipMemcpyAsync(host.data(), dev, 4 * per, hipMemcpyDeviceToHost, op->ctx->stream)
                               : hipSuccess;
    if (e == hipSuccess && rc == GX_OK) e = hipStreamSynchronize(op->ctx->stream);
    if (e != hipSuccess) {
      set_error("DhopDeriv: %s", hipGetErrorString(e));
      rc = GX_EHIP;
      break;
    }
    if (rc != GX_OK) break;
    const int Lx = g.L[0], Ly = g.L[1], Lz = g.L[2], Lxh = Lx / 2;
    for (int64_t h = 0; h < g.H4; h++) {
      int64_t r = h / Lxh;
      const int xh = (int)(h - r * Lxh), y = (int)(r % Ly);
      r /= Ly;
      const int z = (int)(r % Lz), t = (int)(r / Lz);
      const int x = 2 * xh + ((cb + y + z + t) & 1);
      const int64_t site = (((int64_t)t * Lz + z) * Ly + y) * Lx + x;
      for (int mu = 0; mu < 4; mu++)
        std::memcpy(mat + (site * 4 + mu) * 18, host.data() + ((size_t)mu * g.H4 + h) * 18, 18 * sizeof(double));
    }
  }
  (void)hipFree(dev);
  return rc;
}

static int deriv_args(Op *op, double *mat, Field *A, Field *B, bool full) {
  GX_REQUIRE(op && mat && A && B, GX_EINVAL, "DhopDeriv: null argument");
  GX_REQUIRE(same_geom(A->g, op->g) && same_geom(B->g, op->g), GX_EINVAL, "DhopDeriv: field geometry");
  GX_REQUIRE(A->prec == GX_F64 && B->prec == GX_F64, GX_EINVAL, "DhopDeriv: F64 fields");
  GX_REQUIRE((A->subset == GX_FULL) == full && (B->subset == GX_FULL) == full, GX_EINVAL,
             full ? "DhopDeriv: full fields" : "DhopDerivEO/OE: half fields");
  return GX_OK;
}

int gx_DhopDeriv(gx_op op, double *mat, gx_field A, gx_field B, int dag) {
  GX_TRY(deriv_args(O(op), mat, F(A), F(B), true));
  return dhop_deriv(O(op), mat, F(A), F(B), dag, 0, 1);
}
int gx_DhopDerivEO(gx_op op, double *mat, gx_field A, gx_field B, int dag) {
  GX_TRY(deriv_args(O(op), mat, F(A), F(B), false));
  GX_REQUIRE(A->cb == GX_EVEN && B->cb == GX_ODD, GX_ECB, "DhopDerivEO: A even, B odd");
  return dhop_deriv(O(op), mat, F(A), F(B), dag, 0, 0);
}
int gx_DhopDerivOE(gx_op op, double *mat, gx_field A, gx_field B, int dag) {
  GX_TRY(deriv_args(O(op), mat, F(A), F(B), false));
  GX_REQUIRE(A->cb == GX_ODD && B->cb == GX_EVEN, GX_ECB, "DhopDerivOE: A odd, B even");
  return dhop_deriv(O(op), mat, F(A), F(B), dag, 1, 1);
}

static int m5d_common(gx_op op, gx_field psi, gx_field chi, int which, bool with_phi) {
  GX_TRY(check_io(O(op), F(psi), F(chi), false));
  GX_REQUIRE(op->kind == GX_MOBIUS, GX_EINVAL, "M5D family needs a Mobius operator");
  return apply_smat(O(op), F(psi), with_phi ? F(chi) : nullptr, F(chi), which, nullptr);
}
int gx_M5D(gx_op op, gx_field psi, gx_field chi) { return m5d_common(op, psi, chi, Op::S_M5D_OFF, true); }
int gx_M5Ddag(gx_op op, gx_field psi, gx_field chi) { return m5d_common(op, psi, chi, Op::S_M5DDAG_OFF, true); }
int gx_Meooe5D(gx_op op, gx_field psi, gx_field chi) { return m5d_common(op, psi, chi, Op::S_MEOOE5D, false); }
int gx_MeooeDag5D(gx_op op, gx_field psi, gx_field chi) { return m5d_common(op, psi, chi, Op::S_MEOOEDAG5D, false); }

// 4D physical source -> 5D source: tmp[0] = P+ in, tmp[Ls-1] = P- in, then Dminus = 1 - c DW,
// DW = Dhop + (4 - M5).  CayleyFermion5DImplementation.h:101-129 (ImportPhysicalFermionSource, Dminus).
int gx_import_physical_fermion_source(gx_op op_, gx_field in4d, gx_field out5d) {
  Op *op = op_;
  Field *in = in4d, *out = out5d;
  GX_REQUIRE(op && in && out && op->kind == GX_MOBIUS, GX_EINVAL, "import: needs a Mobius operator");
  GX_REQUIRE(in->subset == GX_FULL && out->subset == GX_FULL && in->g.Ls == 1 && same_geom(out->g, op->g) &&
                 in->prec == out->prec && (in->prec == GX_F64 || in->prec == GX_F32),
             GX_EINVAL, "import: needs a full 4D field (Ls=1) and a full 5D field of the operator");
  for (int d = 0; d < 4; d++) GX_REQUIRE(in->g.L[d] == op->g.L[d], GX_EINVAL, "import: 4D dimensions differ");
  const int Ls = op->g.Ls;
  hipStream_t st = op->ctx->stream;
  GX_CHECK_HIP(hipMemsetAsync(out->data, 0, out->bytes, st));
  if (in->prec == GX_F64) {
    GX_TRY(launch_slice_proj<double>(op->g, in->data, 1, 0, out->data, Ls, 0, 0x3, 0, st));
    GX_TRY(launch_slice_proj<double>(op->g, in->data, 1, 0, out->data, Ls, Ls - 1, 0xC, 0, st));
  } else {
    GX_TRY(launch_slice_proj<float>(op->g, in->data, 1, 0, out->data, Ls, 0, 0x3, 0, st));
    GX_TRY(launch_slice_proj<float>(op->g, in->data, 1, 0, out->data, Ls, Ls - 1, 0xC, 0, st));
  }
  Field *t;
  GX_TRY(field_create(op->ctx, op->g, in->prec, GX_FULL, &t));
  int rc = GX_OK;
  for (int cb = 0; cb < 2 && rc == GX_OK; cb++) {
    Field i = half_view(out, 1 - cb), o = half_view(t, cb);
    rc = apply_dhop_half(op, &i, &o, cb, 0);
  }
  const double c = op->cs[0];
  if (rc == GX_OK) rc = blas_axpby(out, 1.0 - c * (4.0 - op->M5), out, -c, t);
  if (hipStreamSynchronize(st) != hipSuccess && rc == GX_OK) rc = GX_EHIP;
  field_destroy(t);
  return rc;
}

// 5D solution -> 4D physical propagator: P- sol[0] + P+ sol[Ls-1].
// CayleyFermion5DImplementation.h:57-67 (ExportPhysicalFermionSolution).
int gx_export_physical_fermion_solution(gx_op op_, gx_field sol5d, gx_field out4d) {
  Op *op = op_;
  Field *sol = sol5d, *out = out4d;
  GX_REQUIRE(op && sol && out && op->kind == GX_MOBIUS, GX_EINVAL, "export: needs a Mobius operator");
  GX_REQUIRE(sol->subset == GX_FULL && out->subset == GX_FULL && out->g.Ls == 1 && same_geom(sol->g, op->g) &&
                 sol->prec == out->prec && (sol->prec == GX_F64 || sol->prec == GX_F32),
             GX_EINVAL, "export: needs a full 5D field of the operator and a full 4D field (Ls=1)");
  for (int d = 0; d < 4; d++) GX_REQUIRE(out->g.L[d] == op->g.L[d], GX_EINVAL, "export: 4D dimensions differ");
  const int Ls = op->g.Ls;
  hipStream_t st = op->ctx->stream;
  if (sol->prec == GX_F64) {
    GX_TRY(launch_slice_proj<double>(op->g, sol->data, Ls, 0, out->data, 1, 0, 0xC, 0, st));
    GX_TRY(launch_slice_proj<double>(op->g, sol->data, Ls, Ls - 1, out->data, 1, 0, 0x3, 1, st));
  } else {
    GX_TRY(launch_slice_proj<float>(op->g, sol->data, Ls, 0, out->data, 1, 0, 0xC, 0, st));
    GX_TRY(launch_slice_proj<float>(op->g, sol->data, Ls, Ls - 1, out->data, 1, 0, 0x3, 1, st));
  }
  return GX_OK;
}

int gx_Mpc(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mpc(O(op), F(in), F(out), 0);
}
int gx_MpcDag(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mpc(O(op), F(in), F(out), 1);
}
int gx_MpcDagMpc(gx_op op, gx_field in, gx_field out) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  return apply_mpcdagmpc(O(op), F(in), F(out));
}
int gx_HermOpAndNorm(gx_op op, gx_field in, gx_field out, double *n1, double *n2) {
  GX_TRY(check_io(O(op), F(in), F(out), false));
  GX_TRY(apply_mpcdagmpc(O(op), F(in), F(out)));
  double im;
  GX_TRY(blas_inner(F(in), F(out), n1, &im));
  return blas_norm2(F(out), n2);
}

}  // extern "C"
