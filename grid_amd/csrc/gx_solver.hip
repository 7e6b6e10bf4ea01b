// Even-odd solvers of libgx.
//
// ConjugateGradient (Grid/algorithms/iterative/ConjugateGradient.h:68-257): Hestenes-Stiefel CG
// on HermOp = MpcDagMpc.  The iteration runs on device scalars: the reduction kernels compute
// a = c/d and b = cp/c in their last block, and every kernel of iteration k checks
// k <= min(k_conv, kmax) before doing anything, so the host enqueues iterations in batches and
// reads the convergence flag without stalling the device between iterations.  The arithmetic is
// the same sequence as Grid's loop: d = Re<p|A p>, a = c/d, r -= a A p, cp = |r|^2, b = cp/c,
// psi += a p, p = b p + r, stop when cp <= tol^2 |src|^2, then the true residual is recomputed.
//
// MixedPrecisionConjugateGradient (ConjugateGradientMixedPrec.h:71-165): FP64 outer residual,
// inner CG in FP32 with adaptive tolerance, FP64 accumulation of the solution, final FP64 CG.
//
// SchurRedBlackDiagMooeeSolve (SchurRedBlack.h:236-291,385-425).
#include <chrono>
#include <climits>
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <vector>

#include "gx_internal.h"

namespace gx {

// One captured batch of solver iterations.  The graph is rebuilt when the buffers it was captured
// with, the batch length or the launch-shape knobs change.
struct BatchGraph {
  hipGraphExec_t exec = nullptr;
  const void *key[8] = {};
  int batch = 0, version = -1;
};

// Enqueues one batch of iterations: body() issues them; with graphs on, the first call captures
// body() into a hipGraph and every call replays it.  The iterations contain kernel launches only
// (no memset or copy nodes); their device-side gate makes replays past convergence no-ops.
template <class F>
static int run_batch(Ctx *ctx, BatchGraph &g, const void *const key[8], int batch, F body) {
  const int ver = tune_version();
  const bool hit = g.exec && g.batch == batch && g.version == ver && std::memcmp(g.key, key, sizeof(g.key)) == 0;
  if (!hit) {
    if (g.exec) {
      (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    GX_CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    if (rc != GX_OK) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    GX_CHECK_HIP(e);
    const hipError_t ie = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    GX_CHECK_HIP(ie);
    std::memcpy(g.key, key, sizeof(g.key));
    g.batch = batch;
    g.version = ver;
  }
  GX_CHECK_HIP(hipGraphLaunch(g.exec, ctx->stream));
  return GX_OK;
}

struct CGWork {
  CGState *dev = nullptr;
  CGState *host = nullptr;  // pinned, 2 slots
  hipEvent_t ev[2] = {};
  BatchGraph graph[3];  // per precision
  double *npart = nullptr;  // the hop tile sums (MrhsWork::npart, owned there)
};

struct MrhsWork {
  MrhsState *dev = nullptr;
  MrhsState *host = nullptr;  // pinned, 2 slots
  double *out_dev = nullptr;  // per-slice results
  double *out_host = nullptr; // pinned mirror
  double *npart = nullptr;    // per-tile slice sums of the fused hop epilogues (Ls >= 2)
  size_t npart_bytes = 0;
  hipEvent_t ev[2] = {};
  BatchGraph graph[3];  // per precision
};

struct MsWork {
  MsState *dev = nullptr;
  MsState *host = nullptr;  // pinned, 2 slots
  hipEvent_t ev[2] = {};
};

struct SolverWork {
  CGWork cg;
  MrhsWork mrhs;
  MsWork ms;  // allocated by the first multishift solve
};

static void free_graphs(BatchGraph *g, int n) {
  for (int i = 0; i < n; i++)
    if (g[i].exec) (void)hipGraphExecDestroy(g[i].exec);
}

void solver_work_free(SolverWork *w) {
  if (!w) return;
  CGWork &c = w->cg;
  MrhsWork &m = w->mrhs;
  free_graphs(c.graph, 3);
  free_graphs(m.graph, 3);
  if (c.dev) (void)hipFree(c.dev);
  if (c.host) (void)hipHostFree(c.host);
  if (m.dev) (void)hipFree(m.dev);
  if (m.out_dev) (void)hipFree(m.out_dev);
  if (m.host) (void)hipHostFree(m.host);
  if (m.out_host) (void)hipHostFree(m.out_host);
  if (m.npart) (void)hipFree(m.npart);
  MsWork &ms = w->ms;
  if (ms.dev) (void)hipFree(ms.dev);
  if (ms.host) (void)hipHostFree(ms.host);
  for (int i = 0; i < 2; i++) {
    if (c.ev[i]) (void)hipEventDestroy(c.ev[i]);
    if (m.ev[i]) (void)hipEventDestroy(m.ev[i]);
    if (ms.ev[i]) (void)hipEventDestroy(ms.ev[i]);
  }
  delete w;
}

// The operator's solver state, allocated on first use: device state, pinned mirrors and events.
static int solver_work(Op *op, SolverWork **out) {
  if (!op->work) {
    SolverWork *w = new SolverWork();
    CGWork &c = w->cg;
    MrhsWork &m = w->mrhs;
    hipError_t e = hipMalloc(&c.dev, sizeof(CGState));
    if (e == hipSuccess) e = hipHostMalloc(&c.host, 2 * sizeof(CGState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&m.dev, sizeof(MrhsState));
    if (e == hipSuccess) e = hipHostMalloc(&m.host, 2 * sizeof(MrhsState), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc(&m.out_dev, sizeof(double) * kMaxRhs);
    if (e == hipSuccess) e = hipHostMalloc(&m.out_host, sizeof(double) * kMaxRhs, hipHostMallocDefault);
    if (e == hipSuccess && op->g.Ls >= 2 && op->g.Ls <= kMaxRhs) {
      m.npart_bytes = hop_norm_bytes(op->g.H4, op->g.Ls);
      e = hipMalloc(&m.npart, m.npart_bytes);
      c.npart = m.npart;
    }
    for (int i = 0; i < 2 && e == hipSuccess; i++) {
      e = hipEventCreateWithFlags(&c.ev[i], hipEventDisableTiming);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&m.ev[i], hipEventDisableTiming);
    }
    if (e != hipSuccess) {
      solver_work_free(w);
      set_error("solver state: %s", hipGetErrorString(e));
      return GX_ENOMEM;
    }
    op->work = w;
  }
  *out = op->work;
  return GX_OK;
}

size_t solver_work_bytes(const SolverWork *w) { return w ? w->mrhs.npart_bytes : 0; }

size_t hop_norm_bytes(int64_t H4, int Ls) { return sizeof(double) * (size_t)hop_norm_tiles(H4, Ls) * Ls; }

// FP64 follows Grid (d = <p, MpcDag Mpc p>, then r -= a MpcDag Mpc p with its norm).  In FP32 and
// FP16 storage, where the tuned hop kernel has the fused epilogues (HopFuse, Ls >= 2), the Mpc
// hop sums |t|^2 (t = Mpc p) per tile, folded to d = |t|^2 and a, and the last MpcDag hop updates
// r -= a MpcDag t and sums |r|^2 per tile without storing MpcDag t, as in the multi-RHS CG.
template <typename T>
static int cg_iteration(Op *op, Field *psi, Field *p, Field *mmp, Field *r, CGState *st, double *npart) {
  Ctx *ctx = op->ctx;
  const int64_t nsites = p->g.half_pad();
  int64_t n1 = 0, n2 = 0;
  Field *t = nullptr;
  if (!std::is_same<T, double>::value && npart) {
    t = op_tmp(op, p->prec, 2);
    if (!t) return GX_ENOMEM;
    HopFuse f1;
    f1.npart = npart;
    f1.nparts = &n1;
    GX_TRY(apply_mpc(op, p, t, 0, st, &f1));
  }
  if (n1 == 0) {
    if (t) GX_TRY(apply_mpc(op, t, mmp, 1, st));
    else GX_TRY(apply_mpcdagmpc(op, p, mmp, st));
    GX_TRY(launch_cg_dot<T>(ctx, nsites, nsites, p->data, mmp->data, st));
  } else {
    GX_TRY(launch_fold_cg(ctx, npart, n1, op->g.Ls, st, 0));
    HopFuse f2;
    f2.npart = npart;
    f2.nparts = &n2;
    f2.r = r->data;
    f2.a = &st->a;  // device address of the state's a (not dereferenced here)
    f2.astride = 0;
    GX_TRY(apply_mpc(op, t, mmp, 1, st, &f2));
  }
  if (n2 > 0) GX_TRY(launch_fold_cg(ctx, npart, n2, op->g.Ls, st, 1));
  else GX_TRY(launch_cg_axpy_norm<T>(ctx, nsites, nsites, r->data, mmp->data, st));
  GX_TRY(launch_cg_update<T>(ctx, nsites, nsites, psi->data, p->data, r->data, st));
  return GX_OK;
}

// Device-resident CG iterations on (psi, p, mmp, r) of precision prec, from the state in w.dev,
// until the state stops them (convergence, or a reliable update with delta > 0) or k passes maxit.
// Batches of iterations are enqueued two ahead so the device never waits for the host; enqueued
// counts the iterations enqueued so far and carries over between calls.  *out: the final state.
static int cg_device_loop(Op *op, int prec, Field *psi, Field *p, Field *mmp, Field *r, CGWork &w, int maxit,
                          int &enqueued, CGState *out) {
  Ctx *ctx = op->ctx;
  const int64_t vol = op->g.half_elems();
  const int batch = vol >= (1 << 20) ? 8 : (vol >= (1 << 16) ? 16 : 32);
  const bool use_graph = tune_get(GX_TUNE_CG_GRAPH, prec) != 0;
  for (int i = 0; i < 4; i++)
    if (!op_tmp(op, prec, i)) return GX_ENOMEM;  // no allocation inside a capture
  auto one_batch = [&]() -> int {
    for (int i = 0; i < batch; i++) {
      GX_TRY(by_prec(prec, [&](auto t) { return cg_iteration<decltype(t)>(op, psi, p, mmp, r, w.dev, w.npart); }));
    }
    return GX_OK;
  };
  auto enqueue = [&](int s) -> int {
    if (use_graph) {
      const void *key[8] = {psi->data, p->data,   mmp->data, r->data, (const void *)(intptr_t)(prec + 1),
                             w.dev,     (const void *)op, (const void *)(uintptr_t)op->epoch};
      GX_TRY(run_batch(ctx, w.graph[prec], key, batch, one_batch));
      enqueued += batch;
    } else {
      for (int i = 0; i < batch && enqueued < maxit; i++, enqueued++) {
        GX_TRY(by_prec(prec, [&](auto t) { return cg_iteration<decltype(t)>(op, psi, p, mmp, r, w.dev, w.npart); }));
      }
    }
    GX_CHECK_HIP(hipMemcpyAsync(&w.host[s], w.dev, sizeof(CGState), hipMemcpyDeviceToHost, ctx->stream));
    GX_CHECK_HIP(hipEventRecord(w.ev[s], ctx->stream));
    return GX_OK;
  };
  int slot = 0;
  GX_TRY(enqueue(0));
  GX_TRY(enqueue(1));
  CGState fin{};
  for (;;) {
    GX_CHECK_HIP(hipEventSynchronize(w.ev[slot]));
    fin = w.host[slot];
    if (fin.k_conv != INT_MAX || fin.k > maxit) break;
    if (enqueued >= maxit) {
      // everything enqueued: wait for the other slot, which holds the last state
      GX_CHECK_HIP(hipEventSynchronize(w.ev[1 - slot]));
      fin = w.host[1 - slot];
      break;
    }
    GX_TRY(enqueue(slot));
    slot = 1 - slot;
  }
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  *out = fin;
  return GX_OK;
}

// Options of one CG run inside the composite solvers (defect correction).  A guess the caller has
// just zeroed skips the guess norm; an initial residual the caller already holds (r0 = src - A psi
// with |r0|^2, per slice for multi-RHS) replaces A psi; inner solves skip the closing true-residual
// check, whose only use is the statistics of a solve (the outer loop recomputes the FP64 residual).
// The iterates are the ones the plain run produces: A 0 = 0 exactly and r0 is the same vector.
struct CGOpts {
  bool zero_guess = false;
  const Field *r0 = nullptr;
  const double *r0_norm = nullptr;
  bool true_residual = true;
};

int cg_run(Op *op, Field *src, Field *psi, double tol, int maxit, gx_solver_stats *out, const CGOpts &opt = CGOpts()) {
  auto t0 = std::chrono::steady_clock::now();
  const int prec = src->prec;
  GX_TRY(op_prepare(op, prec));
  Field *p = op_tmp(op, prec, 4), *mmp = op_tmp(op, prec, 5), *r = op_tmp(op, prec, 6);
  if (!p || !mmp || !r) return GX_ENOMEM;
  p->cb = mmp->cb = r->cb = psi->cb = src->cb;
  gx_solver_stats st{};
  double ssq, guess = 0.0, a;
  GX_TRY(blas_norm2(src, &ssq));
  if (!opt.zero_guess && !opt.r0) GX_TRY(blas_norm2(psi, &guess));
  if (std::isnan(guess) || std::isnan(ssq)) {
    set_error("CG: NaN in source or guess");
    return GX_ENAN;
  }
  Ctx *ctx = op->ctx;
  if (opt.r0) {
    GX_CHECK_HIP(hipMemcpyAsync(r->data, opt.r0->data, r->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GX_CHECK_HIP(hipMemcpyAsync(p->data, opt.r0->data, p->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    a = *opt.r0_norm;
  } else if (guess == 0.0) {
    GX_CHECK_HIP(hipMemcpyAsync(r->data, src->data, src->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GX_CHECK_HIP(hipMemcpyAsync(p->data, src->data, src->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    a = ssq;
  } else {
    GX_TRY(apply_mpcdagmpc(op, psi, mmp));
    GX_TRY(blas_axpby(r, 1.0, src, -1.0, mmp));
    GX_CHECK_HIP(hipMemcpyAsync(p->data, r->data, r->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    GX_TRY(blas_norm2(p, &a));
  }
  double cp = a;
  if (ssq == 0.0) {
    GX_CHECK_HIP(hipMemsetAsync(psi->data, 0, psi->bytes, ctx->stream));
    st.iterations = 1;
    st.converged = 1;
    st.true_residual = 0;
    if (out) *out = st;
    return GX_OK;
  }
  const double rsq = tol * tol * ssq;
  if (cp <= rsq) {
    st.iterations = 0;
    st.converged = 1;
    st.true_residual = std::sqrt(a / ssq);
    st.computed_residual = st.true_residual;
    if (out) *out = st;
    return GX_OK;
  }
  SolverWork *sw = nullptr;
  GX_TRY(solver_work(op, &sw));
  CGWork &w = sw->cg;
  CGState init{};
  init.c = cp;
  init.cp = cp;
  init.rsq = rsq;
  init.k = 1;
  init.k_conv = INT_MAX;
  init.kmax = maxit;
  w.host[0] = init;
  GX_CHECK_HIP(hipMemcpyAsync(w.dev, &w.host[0], sizeof(CGState), hipMemcpyHostToDevice, ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));

  int enqueued = 0;
  CGState fin{};
  GX_TRY(cg_device_loop(op, prec, psi, p, mmp, r, w, maxit, enqueued, &fin));
  const bool converged = fin.k_conv != INT_MAX;
  st.iterations = converged ? fin.k_conv : maxit;
  st.converged = converged;
  st.computed_residual = std::sqrt(fin.cp / ssq);
  // true residual: |MpcDagMpc psi - src| / |src|   (ConjugateGradient.h:201-208)
  if (opt.true_residual) {
    GX_TRY(apply_mpcdagmpc(op, psi, mmp));
    GX_TRY(blas_axpby(p, 1.0, mmp, -1.0, src));
    double rn, sn;
    GX_TRY(blas_norm2(p, &rn));
    GX_TRY(blas_norm2(src, &sn));
    st.true_residual = std::sqrt(rn) / std::sqrt(sn);
  } else {
    st.true_residual = st.computed_residual;
  }
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = st;
  if (!converged) {
    set_error("CG did not converge in %d iterations (residual %g)", maxit, st.computed_residual);
    return GX_ENOCONV;
  }
  return GX_OK;
}

// Outer exit and leg targets of defect correction.  kGridRule is Grid's (ConjugateGradientMixedPrec.h
// :117-151, ConjugateGradientMixedPrecBatched.h:100-213): the loop ends once every |r|^2 <=
// OuterLoopNormMult (100) x stop and leaves the last factor to the FP64 patch-up CG; the inner
// tolerance only ever doubles from InnerTolerance.  kConvergeRule (the multi-RHS red-black solves
// and block CG of this library) applies to FP32 inner solves: one FP64 iteration costs about
// two FP32 ones, so the loop runs until every right-hand side is converged (multiplier 1), and a leg
// whose needed reduction t = kLand sqrt(stop_max / |r|^2_max) is within one FP32 leg's reach
// (t >= kReach32) aims there, the others at the caller's inner tolerance: with a first leg at 1e-6
// the second lands below stop, and the patch-up CG, still run from the fresh residual as in Grid,
// finds it converged and returns at once.  FP16-storage inner solves keep Grid's rule (one such leg
// gains three to four orders, so the legs, not the tolerance, set the count).
enum { kGridRule = 0, kConvergeRule = 1 };
constexpr double kLand = 0.3;
constexpr double kReach32 = 1e-7;

struct OuterPolicy {
  bool grid;
  double mult;
  OuterPolicy(int rule, int inner_prec) : grid(rule == kGridRule || inner_prec != GX_F32), mult(grid ? 100.0 : 1.0) {}
  // the next leg's inner tolerance from the largest residual and stopping norms
  void next_tol(double &inner_tol, double inner_tol0, double nmax, double smax) const {
    if (grid) {
      while (nmax * inner_tol * inner_tol < smax) inner_tol *= 2;
    } else {
      const double t = kLand * std::sqrt(smax / nmax);
      inner_tol = t >= kReach32 ? t : std::max(inner_tol0, t);
    }
  }
};

int mixed_cg_run(Op *op, Field *src_in, Field *sol, double tol, double inner_tol0, int inner_prec, int maxinner,
                 int maxouter, gx_solver_stats *out, int maxpatch = -1, int rule = kGridRule) {
  auto t0 = std::chrono::steady_clock::now();
  gx_solver_stats st{};
  const int cb = src_in->cb;
  sol->cb = cb;
  double src_norm;
  GX_TRY(blas_norm2(src_in, &src_norm));
  const double stop = src_norm * tol * tol;
  Field *tmp_d = op_tmp(op, GX_F64, 7), *src_d = op_tmp(op, GX_F64, 8);
  Field *src_f = op_tmp(op, inner_prec, 7), *sol_f = op_tmp(op, inner_prec, 8);
  if (!tmp_d || !src_d || !src_f || !sol_f) return GX_ENOMEM;
  tmp_d->cb = src_d->cb = src_f->cb = sol_f->cb = cb;
  double inner_tol = inner_tol0;
  int outer = 0;
  const OuterPolicy pol(rule, inner_prec);
  double sol_norm;
  GX_TRY(blas_norm2(sol, &sol_norm));
  bool sol_zero = sol_norm == 0.0;
  Field *res = src_d;         // src - A sol of the latest check
  double norm = 0;
  bool fresh = false;         // res belongs to the current sol
  for (outer = 0; outer < maxouter; outer++) {
    if (sol_zero) {
      res = src_in;           // A 0 = 0
      norm = src_norm;
    } else {
      GX_TRY(apply_mpcdagmpc(op, sol, tmp_d));
      GX_TRY(blas_axpy_norm(src_d, -1.0, tmp_d, src_in, &norm));   // src_d = src - A sol
      res = src_d;
    }
    fresh = true;
    if (norm < pol.mult * stop) break;
    pol.next_tol(inner_tol, inner_tol0, norm, stop);
    GX_TRY(convert_field(src_f, res));
    GX_CHECK_HIP(hipMemsetAsync(sol_f->data, 0, sol_f->bytes, op->ctx->stream));
    gx_solver_stats ist{};
    CGOpts io;
    io.zero_guess = true;
    io.true_residual = false;
    int rc = cg_run(op, src_f, sol_f, inner_tol, maxinner, &ist, io);
    if (rc != GX_OK && rc != GX_ENOCONV) return rc;
    st.inner_iterations += ist.iterations;
    GX_TRY(accumulate_field(sol, sol_f));
    sol_zero = false;
    fresh = false;
  }
  st.outer_iterations = outer;
  gx_solver_stats fst{};
  // the final FP64 solve starts from the residual of the last check when it is current
  CGOpts fo;
  if (fresh) {
    fo.r0 = res;
    fo.r0_norm = &norm;
  }
  const int rc = cg_run(op, src_in, sol, tol, maxpatch > 0 ? maxpatch : maxinner, &fst, fo);
  st.final_iterations = fst.iterations;
  st.iterations = st.inner_iterations + fst.iterations;
  st.true_residual = fst.true_residual;
  st.computed_residual = fst.computed_residual;
  st.converged = fst.converged;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = st;
  return rc;
}

// ConjugateGradientReliableUpdate (ConjugateGradientReliableUpdate.h:77-269): the Krylov vectors
// (psi_f, p, mmp, r) live in the sloppy precision sp while psi accumulates in FP64.  The device
// iteration stops itself when the residual falls below delta times its largest value since the
// last update; the host then folds psi_f into psi, recomputes r = src - A psi in FP64, restarts
// psi_f at zero, forms p = (cp_new / c) p + r, and resumes.  After convergence an FP64 CG from psi
// cleans up (DoFinalCleanup).
//
// setFallbackLinop (ConjugateGradientReliableUpdate.h:72-75,255-259): with fb_prec >= 0 the sloppy
// iteration switches to precision fb_prec once its residual norm cp drops below fb_tol (an absolute
// |r|^2, as in Grid).  Grid switches the operator of its half-precision communications to a
// single-precision one; here the sloppy vectors themselves change storage (FP16 to FP32): p and r
// are promoted exactly, psi_f is folded into psi, and the iteration continues with the same
// scalars.
static int ru_cg_run(Op *op, Field *src, Field *psi, double tol, int maxit, double delta, int sp, int fb_prec,
                     double fb_tol, gx_solver_stats *out) {
  auto t0 = std::chrono::steady_clock::now();
  const bool fallback = fb_prec >= 0 && fb_prec != sp && fb_tol > 0;
  GX_TRY(op_prepare(op, GX_F64));
  GX_TRY(op_prepare(op, sp));
  if (fallback) GX_TRY(op_prepare(op, fb_prec));
  Field *md = op_tmp(op, GX_F64, 7), *rd = op_tmp(op, GX_F64, 8);
  Field *pf = op_tmp(op, sp, 4), *mf = op_tmp(op, sp, 5), *rf = op_tmp(op, sp, 6), *xf = op_tmp(op, sp, 8);
  if (!md || !rd || !pf || !mf || !rf || !xf) return GX_ENOMEM;
  Field *p2 = nullptr, *m2 = nullptr, *r2 = nullptr, *x2 = nullptr;  // fallback-precision vectors
  if (fallback) {
    p2 = op_tmp(op, fb_prec, 4), m2 = op_tmp(op, fb_prec, 5), r2 = op_tmp(op, fb_prec, 6), x2 = op_tmp(op, fb_prec, 8);
    if (!p2 || !m2 || !r2 || !x2) return GX_ENOMEM;
  }
  const int cb = src->cb;
  psi->cb = md->cb = rd->cb = pf->cb = mf->cb = rf->cb = xf->cb = cb;
  Ctx *ctx = op->ctx;
  gx_solver_stats st{};
  double ssq, guess, cp;
  GX_TRY(blas_norm2(src, &ssq));
  GX_TRY(blas_norm2(psi, &guess));
  if (std::isnan(guess) || std::isnan(ssq)) {
    set_error("reliable-update CG: NaN in source or guess");
    return GX_ENAN;
  }
  if (ssq == 0.0) {
    GX_CHECK_HIP(hipMemsetAsync(psi->data, 0, psi->bytes, ctx->stream));
    st.converged = 1;
    if (out) *out = st;
    return GX_OK;
  }
  GX_TRY(apply_mpcdagmpc(op, psi, md));
  GX_TRY(blas_axpy_norm(rd, -1.0, md, src, &cp));  // r = src - A psi
  const double rsq = tol * tol * ssq;
  if (cp <= rsq) {
    st.converged = 1;
    st.true_residual = st.computed_residual = std::sqrt(cp / ssq);
    if (out) *out = st;
    return GX_OK;
  }
  GX_TRY(convert_field(rf, rd));
  GX_CHECK_HIP(hipMemsetAsync(xf->data, 0, xf->bytes, ctx->stream));
  GX_CHECK_HIP(hipMemcpyAsync(pf->data, rf->data, rf->bytes, hipMemcpyDeviceToDevice, ctx->stream));
  SolverWork *sw = nullptr;
  GX_TRY(solver_work(op, &sw));
  CGWork &w = sw->cg;
  CGState init{};
  init.c = init.cp = init.maxres = cp;
  init.rsq = rsq;
  init.delta = delta;
  init.fbtol = fallback ? fb_tol : 0.0;
  init.k = 1;
  init.k_conv = INT_MAX;
  init.kmax = maxit;
  w.host[0] = init;
  GX_CHECK_HIP(hipMemcpyAsync(w.dev, &w.host[0], sizeof(CGState), hipMemcpyHostToDevice, ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  int enqueued = 0, updates = 0, sloppy_its = maxit, cur = sp;
  for (;;) {
    CGState fin{};
    GX_TRY(cg_device_loop(op, cur, xf, pf, mf, rf, w, maxit, enqueued, &fin));
    GX_TRY(convert_field(md, xf));
    GX_TRY(blas_axpby(psi, 1.0, psi, 1.0, md));  // psi += psi_f
    if (fin.k_conv == INT_MAX || !fin.ru) {        // converged, or out of iterations
      if (fin.k_conv != INT_MAX) sloppy_its = fin.k_conv;
      break;
    }
    CGState rs = fin;
    GX_CHECK_HIP(hipMemsetAsync(xf->data, 0, xf->bytes, ctx->stream));
    if (fin.ru == 1) {
      // reliable update (ConjugateGradientReliableUpdate.h:209-231)
      double cpn;
      GX_TRY(apply_mpcdagmpc(op, psi, md));
      GX_TRY(blas_axpy_norm(rd, -1.0, md, src, &cpn));
      GX_TRY(convert_field(rf, rd));
      const double b = cpn / fin.cprev;
      GX_TRY(blas_axpby(pf, b, pf, 1.0, rf));  // p = b p + r (ConjugateGradientReliableUpdate.h:234)
      updates++;
      rs.c = rs.cp = rs.maxres = cpn;
      rs.b = b;
    }
    if (rs.fbtol > 0 && rs.cp < rs.fbtol) {  // switch to the fallback precision (:255-259)
      GX_TRY(convert_field(p2, pf));
      GX_TRY(convert_field(r2, rf));
      x2->cb = m2->cb = cb;
      GX_CHECK_HIP(hipMemsetAsync(x2->data, 0, x2->bytes, ctx->stream));
      pf = p2, mf = m2, rf = r2, xf = x2;
      cur = fb_prec;
      rs.fbtol = 0.0;
    }
    rs.ru = 0;
    rs.k = fin.k_conv + 1;
    rs.k_conv = INT_MAX;
    enqueued = fin.k_conv;
    if (rs.k > maxit) break;
    w.host[0] = rs;
    GX_CHECK_HIP(hipMemcpyAsync(w.dev, &w.host[0], sizeof(CGState), hipMemcpyHostToDevice, ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  }
  gx_solver_stats fst{};
  const int rc = cg_run(op, src, psi, tol, maxit, &fst);  // final FP64 cleanup
  st.inner_iterations = sloppy_its;
  st.outer_iterations = updates;
  st.final_iterations = fst.iterations;
  st.iterations = sloppy_its + fst.iterations;
  st.true_residual = fst.true_residual;
  st.computed_residual = fst.computed_residual;
  st.converged = fst.converged;
  st.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (out) *out = st;
  return rc;
}

// ---------------- multi-RHS CG (BlockConjugateGradient.h:412-545, CGmultiRHS) ----------------
// per-slice reduction that returns host values: mode 1 |x_s|^2, mode 3 z = y - x and |z_s|^2
static int slice_reduce(Op *op, int mode, Field *x, Field *y, Field *z, double *out) {
  SolverWork *sw = nullptr;
  GX_TRY(solver_work(op, &sw));
  MrhsWork &w = sw->mrhs;
  Ctx *ctx = op->ctx;
  const int rc = by_prec(x->prec, [&](auto tag) {
    return launch_mrhs_reduce<decltype(tag)>(ctx, mode, op->g, x->data, y ? y->data : nullptr, z ? z->data : nullptr,
                                             w.dev, w.out_dev);
  });
  GX_TRY(rc);
  GX_CHECK_HIP(hipMemcpyAsync(w.out_host, w.out_dev, sizeof(double) * op->g.Ls, hipMemcpyDeviceToHost, ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  for (int s = 0; s < op->g.Ls; s++) out[s] = w.out_host[s];
  if (z) z->cb = x->cb;
  return GX_OK;
}

// One CGmultiRHS iteration.  d_s = Re<p_s, MpcDag Mpc p_s> as in Grid (BlockConjugateGradient.h:474-478)
// for FP64; the FP32 / FP16-storage inner solves of defect correction take the same quantity as
// |Mpc p_s|^2 from the intermediate t = Mpc p (equal in exact arithmetic since MpcDag is the adjoint
// of Mpc).  There, where the tuned hop kernel has the fused epilogues (HopFuse), the Mpc hop sums
// |t_s|^2 per tile (folded to d_s, a_s), and the last MpcDag hop updates r_s -= a_s (MpcDag t)_s
// and sums |r_s|^2 per tile without storing MpcDag t: the iteration reads and writes no field
// beyond the hops and the x, p update.  Otherwise t, then (r, mmp) are reduced by their own passes.
template <typename T>
static int mrhs_iteration(Op *op, Field *psi, Field *p, Field *mmp, Field *r, MrhsState *st, double *npart) {
  Ctx *ctx = op->ctx;
  if (std::is_same<T, double>::value || !npart) {
    GX_TRY(apply_mpcdagmpc(op, p, mmp, st));
    if (std::is_same<T, double>::value) {
      GX_TRY(launch_mrhs_reduce<T>(ctx, 0, op->g, p->data, mmp->data, nullptr, st, nullptr));
    } else {
      Field *t = op_tmp(op, p->prec, 2);
      if (!t) return GX_ENOMEM;
      GX_TRY(launch_mrhs_reduce<T>(ctx, 4, op->g, t->data, nullptr, nullptr, st, nullptr));
    }
    GX_TRY(launch_mrhs_reduce<T>(ctx, 2, op->g, r->data, mmp->data, nullptr, st, nullptr));
  } else {
    Field *t = op_tmp(op, p->prec, 2);
    if (!t) return GX_ENOMEM;
    int64_t n1 = 0, n2 = 0;
    HopFuse f1;
    f1.npart = npart;
    f1.nparts = &n1;
    GX_TRY(apply_mpc(op, p, t, 0, st, &f1));
    if (n1 > 0) GX_TRY(launch_fold_norms(ctx, npart, n1, op->g.Ls, st, 4));
    else GX_TRY(launch_mrhs_reduce<T>(ctx, 4, op->g, t->data, nullptr, nullptr, st, nullptr));
    HopFuse f2;
    f2.npart = npart;
    f2.nparts = &n2;
    f2.r = r->data;
    f2.a = st->a;  // device address of the state's a_s (not dereferenced here)
    GX_TRY(apply_mpc(op, t, mmp, 1, st, &f2));
    if (n2 > 0) GX_TRY(launch_fold_norms(ctx, npart, n2, op->g.Ls, st, 2));
    else GX_TRY(launch_mrhs_reduce<T>(ctx, 2, op->g, r->data, mmp->data, nullptr, st, nullptr));
  }
  GX_TRY(launch_mrhs_update<T>(ctx, op->g, psi->data, p->data, r->data, st));
  return GX_OK;
}

// CGmultiRHS on HermOp = MpcDagMpc with right-hand side k in slice s = k.  st: Ls entries or null.
int mrhs_cg_run(Op *op, Field *src, Field *psi, double tol, int maxit, gx_solver_stats *st,
                const CGOpts &opt = CGOpts()) {
  auto t0 = std::chrono::steady_clock::now();
  const int Ls = op->g.Ls, prec = src->prec;
  if (Ls > kMaxRhs) {
    set_error("multi-RHS CG: %d right-hand sides (at most %d)", Ls, kMaxRhs);
    return GX_EINVAL;
  }
  GX_TRY(op_prepare(op, prec));
  Field *p = op_tmp(op, prec, 4), *mmp = op_tmp(op, prec, 5), *r = op_tmp(op, prec, 6);
  if (!p || !mmp || !r) return GX_ENOMEM;
  p->cb = mmp->cb = r->cb = psi->cb = src->cb;
  std::vector<double> ssq(Ls), rr(Ls), gn(Ls, 0.0);
  GX_TRY(slice_reduce(op, 1, src, nullptr, nullptr, ssq.data()));
  for (int s = 0; s < Ls; s++)
    if (std::isnan(ssq[s])) {
      set_error("block CG: NaN in source slice %d", s);
      return GX_ENAN;
    }
  Ctx *ctx = op->ctx;
  bool zero = opt.zero_guess;
  if (!zero && !opt.r0) {
    GX_TRY(slice_reduce(op, 1, psi, nullptr, nullptr, gn.data()));
    zero = true;
    for (int s = 0; s < Ls; s++) zero = zero && gn[s] == 0.0;
  }
  // a caller's residual (true, of this guess) that already meets tol on every slice: no
  // iteration, and its norms are the true residual
  bool r0_done = opt.r0 != nullptr;
  if (opt.r0)
    for (int s = 0; s < Ls; s++) {
      rr[s] = opt.r0_norm[s];
      r0_done = r0_done && rr[s] < tol * tol * (ssq[s] > 0 ? ssq[s] : 1.0);
    }
  if (opt.r0) {
    if (!r0_done)
      GX_CHECK_HIP(hipMemcpyAsync(r->data, opt.r0->data, r->bytes, hipMemcpyDeviceToDevice, ctx->stream));
  } else if (zero) {
    // A 0 = 0: R = src
    GX_CHECK_HIP(hipMemcpyAsync(r->data, src->data, r->bytes, hipMemcpyDeviceToDevice, ctx->stream));
    rr = ssq;
  } else {
    // initial search direction from the guess: AP = A psi, R = src - AP, P = R
    GX_TRY(apply_mpcdagmpc(op, psi, mmp));
    GX_TRY(slice_reduce(op, 3, mmp, src, r, rr.data()));
  }
  for (int s = 0; s < Ls; s++)
    if (std::isnan(rr[s])) {
      set_error("block CG: NaN in guess slice %d", s);
      return GX_ENAN;
    }
  if (!r0_done) GX_CHECK_HIP(hipMemcpyAsync(p->data, r->data, r->bytes, hipMemcpyDeviceToDevice, ctx->stream));
  SolverWork *sw = nullptr;
  GX_TRY(solver_work(op, &sw));
  MrhsWork &w = sw->mrhs;
  MrhsState init{};
  double mx = 0;
  for (int s = 0; s < Ls; s++) {
    init.c[s] = rr[s];
    init.cp[s] = rr[s];
    init.ssq[s] = ssq[s] > 0 ? ssq[s] : 1.0;
    mx = std::max(mx, rr[s] / init.ssq[s]);
  }
  init.tol2 = tol * tol;
  init.k = 1;
  init.k_conv = INT_MAX;
  init.kmax = maxit;
  int iterations = 0;
  bool converged = false;
  MrhsState fin{};
  if (mx < init.tol2) {
    converged = true;
    fin = init;
  } else {
    w.host[0] = init;
    GX_CHECK_HIP(hipMemcpyAsync(w.dev, &w.host[0], sizeof(MrhsState), hipMemcpyHostToDevice, ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    const int64_t vol = op->g.half_elems();
    // iterations per enqueued batch: one batch at most runs gated past convergence
    const int batch = vol >= (1 << 24) ? 4 : (vol >= (1 << 20) ? 8 : (vol >= (1 << 16) ? 16 : 32));
    int enqueued = 0, slot = 0;
    const bool use_graph = tune_get(GX_TUNE_CG_GRAPH, prec) != 0;
    for (int i = 0; i < 4; i++)
      if (!op_tmp(op, prec, i)) return GX_ENOMEM;  // no allocation inside a capture
    auto one_batch = [&]() -> int {
      for (int i = 0; i < batch; i++) {
        int rc = by_prec(prec, [&](auto tag) { return mrhs_iteration<decltype(tag)>(op, psi, p, mmp, r, w.dev, w.npart); });
        GX_TRY(rc);
      }
      return GX_OK;
    };
    auto enqueue = [&](int sl) -> int {
      if (use_graph) {
        const void *key[8] = {psi->data, p->data,   mmp->data, r->data, (const void *)(intptr_t)(prec + 1),
                             w.dev,     (const void *)op, (const void *)(uintptr_t)op->epoch};
        GX_TRY(run_batch(ctx, w.graph[prec], key, batch, one_batch));
        enqueued += batch;
      } else {
        for (int i = 0; i < batch && enqueued < maxit; i++, enqueued++) {
          int rc = by_prec(prec, [&](auto tag) { return mrhs_iteration<decltype(tag)>(op, psi, p, mmp, r, w.dev, w.npart); });
          GX_TRY(rc);
        }
      }
      GX_CHECK_HIP(hipMemcpyAsync(&w.host[sl], w.dev, sizeof(MrhsState), hipMemcpyDeviceToHost, ctx->stream));
      GX_CHECK_HIP(hipEventRecord(w.ev[sl], ctx->stream));
      return GX_OK;
    };
    GX_TRY(enqueue(0));
    GX_TRY(enqueue(1));
    for (;;) {
      GX_CHECK_HIP(hipEventSynchronize(w.ev[slot]));
      fin = w.host[slot];
      if (fin.k_conv != INT_MAX || fin.k > maxit) break;
      if (enqueued >= maxit) {
        GX_CHECK_HIP(hipEventSynchronize(w.ev[1 - slot]));
        fin = w.host[1 - slot];
        break;
      }
      GX_TRY(enqueue(slot));
      slot = 1 - slot;
    }
    GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    converged = fin.k_conv != INT_MAX;
    iterations = converged ? fin.k_conv : maxit;
  }
  // true residual per right-hand side (inner solves report the recursive one)
  std::vector<double> res(Ls);
  if (r0_done) {
    res = rr;
  } else if (opt.true_residual) {
    GX_TRY(apply_mpcdagmpc(op, psi, mmp));
    GX_TRY(slice_reduce(op, 3, src, mmp, p, res.data()));
  } else {
    for (int s = 0; s < Ls; s++) res[s] = fin.cp[s];
  }
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st)
    for (int s = 0; s < Ls; s++) {
      st[s] = gx_solver_stats{};
      st[s].iterations = iterations;
      st[s].converged = converged;
      st[s].true_residual = std::sqrt(res[s] / (ssq[s] > 0 ? ssq[s] : 1.0));
      st[s].computed_residual = std::sqrt(fin.cp[s] / (ssq[s] > 0 ? ssq[s] : 1.0));
      st[s].seconds = secs;
    }
  if (!converged) {
    set_error("block CG did not converge in %d iterations", maxit);
    return GX_ENOCONV;
  }
  return GX_OK;
}

// Batched defect correction (ConjugateGradientMixedPrecBatched.h:100-213) with the inner solves of
// all right-hand sides done together by one multi-RHS CG in FP32 or FP16 storage (inner_prec), so
// they share every link load.  rule: the outer policy (OuterPolicy), on the largest residual and
// stopping norm over the right-hand sides.
int mrhs_mixed_run(Op *op, Field *src_in, Field *sol, double tol, double inner_tol0, int inner_prec, int maxinner,
                   int maxouter, gx_solver_stats *st, int maxpatch = -1, int rule = kGridRule) {
  auto t0 = std::chrono::steady_clock::now();
  const int Ls = op->g.Ls, cb = src_in->cb;
  sol->cb = cb;
  std::vector<double> src_norm(Ls), stop(Ls), norm(Ls);
  GX_TRY(slice_reduce(op, 1, src_in, nullptr, nullptr, src_norm.data()));
  for (int s = 0; s < Ls; s++) stop[s] = src_norm[s] * tol * tol;
  Field *tmp_d = op_tmp(op, GX_F64, 7), *src_d = op_tmp(op, GX_F64, 8);
  Field *src_f = op_tmp(op, inner_prec, 7), *sol_f = op_tmp(op, inner_prec, 8);
  if (!tmp_d || !src_d || !src_f || !sol_f) return GX_ENOMEM;
  tmp_d->cb = src_d->cb = src_f->cb = sol_f->cb = cb;
  int outer = 0, inner_total = 0;
  std::vector<double> sn(Ls);
  GX_TRY(slice_reduce(op, 1, sol, nullptr, nullptr, sn.data()));
  bool sol_zero = true;
  for (int s = 0; s < Ls; s++) sol_zero = sol_zero && sn[s] == 0.0;
  Field *res = src_d;   // src - A sol of the latest check
  bool fresh = false;   // res belongs to the current sol
  const OuterPolicy pol(rule, inner_prec);
  double inner_tol = inner_tol0;
  for (outer = 0; outer < maxouter; outer++) {
    if (sol_zero) {
      res = src_in;     // A 0 = 0
      norm = src_norm;
    } else {
      GX_TRY(apply_mpcdagmpc(op, sol, tmp_d));
      GX_TRY(slice_reduce(op, 3, tmp_d, src_in, src_d, norm.data()));   // src_d = src - A sol
      res = src_d;
    }
    fresh = true;
    bool all = true;
    double nmax = 0, smax = 0;
    for (int s = 0; s < Ls; s++) {
      if (norm[s] > pol.mult * stop[s]) all = false;
      nmax = std::max(nmax, norm[s]);
      smax = std::max(smax, stop[s]);
    }
    if (all) break;
    pol.next_tol(inner_tol, inner_tol0, nmax, smax);
    GX_TRY(convert_field(src_f, res));
    GX_CHECK_HIP(hipMemsetAsync(sol_f->data, 0, sol_f->bytes, op->ctx->stream));
    std::vector<gx_solver_stats> ist(Ls);
    CGOpts io;
    io.zero_guess = true;
    io.true_residual = false;
    int rc = mrhs_cg_run(op, src_f, sol_f, inner_tol, maxinner, ist.data(), io);
    if (rc != GX_OK && rc != GX_ENOCONV) return rc;
    inner_total += ist[0].iterations;
    GX_TRY(accumulate_field(sol, sol_f));
    sol_zero = false;
    fresh = false;
  }
  std::vector<gx_solver_stats> fst(Ls);
  // the final FP64 solve starts from the residual of the last check when it is current
  CGOpts fo;
  if (fresh) {
    fo.r0 = res;
    fo.r0_norm = norm.data();
  }
  const int rc = mrhs_cg_run(op, src_in, sol, tol, maxpatch > 0 ? maxpatch : maxinner, fst.data(), fo);
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st)
    for (int s = 0; s < Ls; s++) {
      st[s] = fst[s];
      st[s].inner_iterations = inner_total;
      st[s].outer_iterations = outer;
      st[s].final_iterations = fst[s].iterations;
      st[s].iterations = inner_total + fst[s].iterations;
      st[s].seconds = secs;
    }
  return rc;
}

// MixedPrecisionConjugateGradientBatched (ConjugateGradientMixedPrecBatched.h:77-202) over n
// sources of one checkerboard.  On a single-RHS Wilson operator the sources are packed into the
// slices of one multi-RHS field of a clone that reads the operator's links, and the batch runs as
// mrhs_mixed_run: one FP64 residual per source per outer iteration, one inner tolerance for the
// batch from the largest residual (as Grid's updateResidual), inner solves of all sources in one
// multi-RHS CG that loads every link once for the whole batch, and the patch-up CG.  Other
// operators (Mobius, or more sources than a clone holds) run Grid's per-source defect correction.
static int mixed_batched_run(Op *op, int n, Field **src, Field **sol, double tol, double inner_tol, int inner_prec,
                             int maxinner, int maxouter, int maxpatch, gx_solver_stats *st) {
  const bool pack = op->kind == GX_WILSON && op->g.Ls == 1 && n >= 2 && n <= kMaxLs;
  if (!pack) {
    int rc = GX_OK;
    for (int k = 0; k < n; k++) {
      const int r = mixed_cg_run(op, src[k], sol[k], tol, inner_tol, inner_prec, maxinner, maxouter, st ? &st[k] : nullptr,
                                 maxpatch);
      if (r != GX_OK && r != GX_ENOCONV) return r;
      if (r == GX_ENOCONV) rc = r;
    }
    return rc;
  }
  Op *base = op->link_owner ? op->link_owner : op;
  gx_op cl = nullptr;
  const double mass = op->slice_mass.empty() ? op->mass : op->slice_mass[0];  // Ls = 1
  GX_TRY(gx_wilson_clone(static_cast<gx_op>(base), static_cast<gx_ctx>(op->ctx), n, mass, &cl));
  Op *B = static_cast<Op *>(cl);
  Field *ps = nullptr, *px = nullptr;
  int rc = field_create(op->ctx, B->g, GX_F64, src[0]->cb, &ps);
  if (rc == GX_OK) rc = field_create(op->ctx, B->g, GX_F64, src[0]->cb, &px);
  for (int k = 0; k < n && rc == GX_OK; k++) {
    rc = slice_transfer(src[k], ps, k, false);
    if (rc == GX_OK) rc = slice_transfer(sol[k], px, k, false);
  }
  std::vector<gx_solver_stats> bst(n);
  if (rc == GX_OK) {
    ps->cb = px->cb = src[0]->cb;
    rc = mrhs_mixed_run(B, ps, px, tol, inner_tol, inner_prec, maxinner, maxouter, bst.data(), maxpatch);
    if (rc == GX_OK || rc == GX_ENOCONV) {
      const int solve_rc = rc;
      for (int k = 0; k < n && rc == solve_rc; k++)
        if (slice_transfer(sol[k], px, k, true) != GX_OK) rc = GX_EHIP;
    }
  }
  if (hipStreamSynchronize(op->ctx->stream) != hipSuccess && (rc == GX_OK || rc == GX_ENOCONV)) rc = GX_EHIP;
  field_destroy(ps);
  field_destroy(px);
  gx_op_destroy(cl);
  if (st)
    for (int k = 0; k < n; k++) st[k] = bst[k];
  return rc;
}

// ---------------- BlockCGrQ (BlockConjugateGradient.h:223-410) ----------------
// Small dense complex matrices (n = Ls <= 16, row-major) for the block solver's host algebra.
typedef std::complex<double> cd;
struct CMat {
  int n = 0;
  std::vector<cd> a;
  explicit CMat(int n_ = 0, bool ident = false) : n(n_), a((size_t)n_ * n_) {
    if (ident)
      for (int i = 0; i < n; i++) a[(size_t)i * n + i] = 1.0;
  }
  cd &operator()(int i, int j) { return a[(size_t)i * n + j]; }
  cd operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

static CMat cm_mul(const CMat &x, const CMat &y) {
  CMat z(x.n);
  for (int i = 0; i < x.n; i++)
    for (int k = 0; k < x.n; k++) {
      const cd v = x(i, k);
      for (int j = 0; j < x.n; j++) z(i, j) += v * y(k, j);
    }
  return z;
}

static CMat cm_adj(const CMat &x) {
  CMat z(x.n);
  for (int i = 0; i < x.n; i++)
    for (int j = 0; j < x.n; j++) z(i, j) = std::conj(x(j, i));
  return z;
}

// Gauss-Jordan inverse with partial pivoting; false if a pivot is zero or not finite.
static bool cm_inv(const CMat &x, CMat &out) {
  const int n = x.n;
  CMat w = x;
  out = CMat(n, true);
  for (int c = 0; c < n; c++) {
    int piv = c;
    for (int r = c + 1; r < n; r++)
      if (std::abs(w(r, c)) > std::abs(w(piv, c))) piv = r;
    const double pm = std::abs(w(piv, c));
    if (!(pm > 0.0) || !std::isfinite(pm)) return false;
    if (piv != c)
      for (int j = 0; j < n; j++) {
        std::swap(w(piv, j), w(c, j));
        std::swap(out(piv, j), out(c, j));
      }
    const cd inv = 1.0 / w(c, c);
    for (int j = 0; j < n; j++) {
      w(c, j) *= inv;
      out(c, j) *= inv;
    }
    for (int r = 0; r < n; r++)
      if (r != c && w(r, c) != 0.0) {
        const cd f = w(r, c);
        for (int j = 0; j < n; j++) {
          w(r, j) -= f * w(c, j);
          out(r, j) -= f * out(c, j);
        }
      }
  }
  return true;
}

// Cholesky m = L L^dag of a Hermitian positive-definite m (Eigen's llt().matrixL()); false if m is
// not positive definite to working precision.
static bool cm_chol(const CMat &m, CMat &L) {
  const int n = m.n;
  L = CMat(n);
  for (int j = 0; j < n; j++) {
    double d = m(j, j).real();
    for (int k = 0; k < j; k++) d -= std::norm(L(j, k));
    if (!(d > 0.0) || !std::isfinite(d)) return false;
    const double ljj = std::sqrt(d);
    L(j, j) = ljj;
    for (int i = j + 1; i < n; i++) {
      cd v = m(i, j);
      for (int k = 0; k < j; k++) v -= L(i, k) * std::conj(L(j, k));
      L(i, j) = v / ljj;
    }
  }
  return true;
}

constexpr int kIpmParts = 2048;  // workgroups of the slice inner products (8 per CU)

struct BlockWork {
  Ctx *ctx = nullptr;
  void *partial = nullptr;   // workgroup partials of the slice inner products
  double2 *mat_dev = nullptr;   // result matrix + 4 coefficient matrices, kMaxLs^2 each
  double2 *mat_host = nullptr;  // pinned mirror
  int nparts = kIpmParts;
  ~BlockWork() {
    if (partial) (void)hipFree(partial);
    if (mat_dev) (void)hipFree(mat_dev);
    if (mat_host) (void)hipHostFree(mat_host);
  }
  int init(Ctx *c) {
    ctx = c;
    const size_t m = (size_t)kMaxLs * kMaxLs * sizeof(double2);
    hipError_t e = hipMalloc(&partial, (size_t)nparts * m);
    if (e == hipSuccess) e = hipMalloc(&mat_dev, 5 * m);
    if (e == hipSuccess) e = hipHostMalloc(&mat_host, 5 * m, hipHostMallocDefault);
    if (e != hipSuccess) {
      set_error("block solver workspace: %s", hipGetErrorString(e));
      return GX_ENOMEM;
    }
    return GX_OK;
  }
  // m(s, t) = <x_s, y_t> (sliceInnerProductMatrix); synchronises the stream
  int ipm(const Geom &g, int nhalf, Field *x, Field *y, CMat &m) {
    const int n = g.Ls;
    GX_TRY(launch_slice_ipm(ctx, g, nhalf, x->data, y->data, partial, nparts, mat_dev));
    GX_CHECK_HIP(hipMemcpyAsync(mat_host, mat_dev, sizeof(double2) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    m = CMat(n);
    for (int i = 0; i < n * n; i++) m.a[i] = cd(mat_host[i].x, mat_host[i].y);
    return GX_OK;
  }
  // InnerProductMatrix over n fields of nelem stored complex numbers; synchronises the stream
  int ipm_vec(int n, int64_t nelem, const VecPtrs &x, const VecPtrs &y, CMat &m) {
    GX_TRY(launch_vec_ipm(ctx, nelem, n, x, y, partial, nparts, mat_dev));
    GX_CHECK_HIP(hipMemcpyAsync(mat_host, mat_dev, sizeof(double2) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    m = CMat(n);
    for (int i = 0; i < n * n; i++) m.a[i] = cd(mat_host[i].x, mat_host[i].y);
    return GX_OK;
  }
  // MaddMatrix over n fields (coefficient slots as in madd)
  int madd_vec(int n, int64_t nelem, const VecPtrs &r, const CMat &a, const VecPtrs &x, const VecPtrs &y, double scale,
               int slot) {
    double2 *h = mat_host + (size_t)slot * kMaxLs * kMaxLs, *d = mat_dev + (size_t)slot * kMaxLs * kMaxLs;
    for (int i = 0; i < n * n; i++) h[i] = make_double2(a.a[i].real(), a.a[i].imag());
    GX_CHECK_HIP(hipMemcpyAsync(d, h, sizeof(double2) * n * n, hipMemcpyHostToDevice, ctx->stream));
    return launch_vec_madd(ctx, nelem, n, d, scale, x, y, r);
  }
  // r_t = y_t + scale sum_s x_s a(s, t) (sliceMaddMatrix; y = null: sliceMulMatrix) with coefficient
  // slot 1..4: a slot's pinned copy is rewritten only after a later ipm() has synchronised the
  // stream past its use.
  int madd(const Geom &g, int nhalf, Field *r, const CMat &a, Field *x, Field *y, double scale, int slot) {
    const int n = g.Ls;
    double2 *h = mat_host + (size_t)slot * kMaxLs * kMaxLs, *d = mat_dev + (size_t)slot * kMaxLs * kMaxLs;
    for (int i = 0; i < n * n; i++) h[i] = make_double2(a.a[i].real(), a.a[i].imag());
    GX_CHECK_HIP(hipMemcpyAsync(d, h, sizeof(double2) * n * n, hipMemcpyHostToDevice, ctx->stream));
    GX_TRY(launch_slice_madd(ctx, g, nhalf, d, scale, x->data, y ? y->data : nullptr, r->data));
    r->cb = x->cb;
    return GX_OK;
  }
};

// The block of Nblock right-hand sides BlockCGrQ works on, in one of two forms:
//   SliceBlock  the fifth-dimension slices of one field (Orthog = 0, BlockCGrQsolve,
//               BlockConjugateGradient.h:223-410); A = MpcDagMpc of a multi-RHS Wilson operator;
//   VecBlock    a list of fields (BlockCGrQsolveVec, :554-741); A = MpcDagMpc on half fields or
//               Mdag M on full fields of any operator, applied field by field.
// Block vectors are named by id: B (sources), X (solutions), R, Q, D, Z (work).
enum { BV_B = 0, BV_X, BV_R, BV_Q, BV_D, BV_Z, BV_COUNT };

struct SliceBlock {
  Op *op;
  BlockWork *bw;
  Field *v[BV_COUNT] = {};
  int n() const { return op->g.Ls; }
  int norms(int id, double *out) { return slice_reduce(op, 1, v[id], nullptr, nullptr, out); }
  int apply(int in, int out) { return apply_mpcdagmpc(op, v[in], v[out]); }
  int sub(int out, int x, int y) { return blas_axpby(v[out], 1.0, v[x], -1.0, v[y]); }  // out = x - y
  int copy(int out, int in) {
    GX_CHECK_HIP(hipMemcpyAsync(v[out]->data, v[in]->data, v[in]->bytes, hipMemcpyDeviceToDevice, op->ctx->stream));
    v[out]->cb = v[in]->cb;
    return GX_OK;
  }
  int ipm(int x, int y, CMat &m) { return bw->ipm(op->g, 1, v[x], v[y], m); }
  int madd(int r, const CMat &a, int x, int y, double scale, int slot) {
    return bw->madd(op->g, 1, v[r], a, v[x], y < 0 ? nullptr : v[y], scale, slot);
  }
  // res[b] = |A X_b - B_b|^2 (R is overwritten)
  int residuals(double *res) {
    GX_TRY(apply_mpcdagmpc(op, v[BV_X], v[BV_Z]));
    return slice_reduce(op, 3, v[BV_Z], v[BV_B], v[BV_R], res);
  }
};

struct VecBlock {
  Op *op;
  BlockWork *bw;
  int herm;                  // GX_HERM_SCHUR or GX_HERM_MDAGM
  int nb = 0;
  Field *full_tmp = nullptr;  // Mdag M: M x
  std::vector<Field *> v[BV_COUNT];
  int n() const { return nb; }
  int apply_one(Field *in, Field *out) {
    if (herm == GX_HERM_SCHUR) return apply_mpcdagmpc(op, in, out);
    GX_TRY(apply_full_m(op, in, full_tmp, 0));
    return apply_full_m(op, full_tmp, out, 1);
  }
  int norms(int id, double *out) {
    for (int b = 0; b < nb; b++) GX_TRY(blas_norm2(v[id][b], &out[b]));
    return GX_OK;
  }
  int apply(int in, int out) {
    for (int b = 0; b < nb; b++) GX_TRY(apply_one(v[in][b], v[out][b]));
    return GX_OK;
  }
  int sub(int out, int x, int y) {
    for (int b = 0; b < nb; b++) GX_TRY(blas_axpby(v[out][b], 1.0, v[x][b], -1.0, v[y][b]));
    return GX_OK;
  }
  int copy(int out, int in) {
    for (int b = 0; b < nb; b++) {
      GX_CHECK_HIP(hipMemcpyAsync(v[out][b]->data, v[in][b]->data, v[in][b]->bytes, hipMemcpyDeviceToDevice,
                                  op->ctx->stream));
      v[out][b]->cb = v[in][b]->cb;
    }
    return GX_OK;
  }
  VecPtrs ptrs(int id) const {
    VecPtrs p{};
    if (id >= 0)
      for (int b = 0; b < nb; b++) p.p[b] = v[id][b]->data;
    return p;
  }
  int ipm(int x, int y, CMat &m) { return bw->ipm_vec(nb, v[x][0]->stored(), ptrs(x), ptrs(y), m); }
  int madd(int r, const CMat &a, int x, int y, double scale, int slot) {
    GX_TRY(bw->madd_vec(nb, v[x][0]->stored(), ptrs(r), a, ptrs(x), ptrs(y), scale, slot));
    for (int b = 0; b < nb; b++) v[r][b]->cb = v[x][b]->cb;
    return GX_OK;
  }
  int residuals(double *res) {
    for (int b = 0; b < nb; b++) {
      GX_TRY(apply_one(v[BV_X][b], v[BV_Z][b]));
      GX_TRY(blas_axpy_norm(v[BV_R][b], -1.0, v[BV_B][b], v[BV_Z][b], &res[b]));
    }
    return GX_OK;
  }
};

// ThinQRfact (BlockConjugateGradient.h:132-190): m_rr = R^dag R (made manifestly Hermitian),
// L = chol(m_rr), C = L^dag, Q = R C^-1.
template <class Blk>
static int thin_qr(Blk &blk, int Q, int R, CMat &C, int slot) {
  CMat m, L, Cinv;
  GX_TRY(blk.ipm(R, R, m));
  CMat h = m;
  for (int i = 0; i < m.n; i++)
    for (int j = 0; j < m.n; j++) h(i, j) = 0.5 * (m(i, j) + std::conj(m(j, i)));
  if (!cm_chol(h, L)) {
    set_error("BlockCGrQ: the residual block is rank deficient (Cholesky of R^dag R failed)");
    return GX_ENOCONV;
  }
  C = cm_adj(L);
  if (!cm_inv(C, Cinv)) {
    set_error("BlockCGrQ: singular triangular factor");
    return GX_ENOCONV;
  }
  return blk.madd(Q, Cinv, R, -1, 1.0, slot);  // MulMatrix(Q, Cinv, R)
}

// BlockCGrQ (Birk's thesis after Dubrulle 2001; BlockConjugateGradient.h:223-410 and the Vec form
// :554-741):
//   QC = R = B - A X (thin QR), D = Q;  for k:  Z = A D,  M = (D^dag Z)^-1,  X += D (M C),
//   QS = Q - Z M (thin QR),  D = Q + D S^dag,  C = S C;  stop when max_b (C^dag C)_bb / |B_b|^2 < tol^2.
// Every right-hand side must be nonzero (the block must have full rank).
template <class Blk>
static int block_cgrq_core(Blk &blk, double tol, int maxit, gx_solver_stats *st) {
  const auto t0 = std::chrono::steady_clock::now();
  const int n = blk.n();
  std::vector<double> ssq(n), xn(n), res(n);
  GX_TRY(blk.norms(BV_B, ssq.data()));
  GX_TRY(blk.norms(BV_X, xn.data()));
  for (int b = 0; b < n; b++) {
    if (std::isnan(ssq[b]) || std::isnan(xn[b])) {
      set_error("BlockCGrQ: NaN in source or guess (right-hand side %d)", b);
      return GX_ENAN;
    }
    if (ssq[b] == 0.0) {
      set_error("BlockCGrQ: right-hand side %d is zero (the block must have full rank)", b);
      return GX_EINVAL;
    }
  }
  GX_TRY(blk.apply(BV_X, BV_Z));
  GX_TRY(blk.sub(BV_R, BV_B, BV_Z));  // R = B - A X
  CMat C, S, M, DZ, rr;
  GX_TRY(thin_qr(blk, BV_Q, BV_R, C, 3));
  GX_TRY(blk.copy(BV_D, BV_Q));
  int rc = GX_ENOCONV, k;
  for (k = 1; k <= maxit; k++) {
    GX_TRY(blk.apply(BV_D, BV_Z));  // Z = A D
    GX_TRY(blk.ipm(BV_D, BV_Z, DZ));
    if (!cm_inv(DZ, M)) {
      set_error("BlockCGrQ: D^dag A D is singular at iteration %d", k);
      break;
    }
    GX_TRY(blk.madd(BV_X, cm_mul(M, C), BV_D, BV_X, 1.0, 1));  // X = X + D M C
    GX_TRY(blk.madd(BV_R, M, BV_Z, BV_Q, -1.0, 2));            // R = Q - Z M
    const int qr = thin_qr(blk, BV_Q, BV_R, S, 3);             // Q S = R
    if (qr == GX_ENOCONV) break;
    GX_TRY(qr);
    GX_TRY(blk.madd(BV_D, cm_adj(S), BV_D, BV_Q, 1.0, 4));     // D = Q + D S^dag
    C = cm_mul(S, C);
    rr = cm_mul(cm_adj(C), C);
    double mx = 0;
    for (int b = 0; b < n; b++) mx = std::max(mx, rr(b, b).real() / ssq[b]);
    if (mx < tol * tol) {
      rc = GX_OK;
      break;
    }
  }
  GX_TRY(blk.residuals(res.data()));  // true residuals |B_b - A X_b| / |B_b|
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (st)
    for (int b = 0; b < n; b++) {
      st[b] = gx_solver_stats{};
      st[b].iterations = std::min(k, maxit);
      st[b].converged = rc == GX_OK;
      st[b].true_residual = std::sqrt(res[b] / ssq[b]);
      st[b].computed_residual = rr.n ? std::sqrt(std::max(0.0, rr(b, b).real()) / ssq[b]) : 0.0;
      st[b].seconds = secs;
    }
  return rc;
}

// BlockCGrQ over the fifth-dimension slices of a multi-RHS Wilson operator (Nblock = Ls; Grid's
// code pins Nblock to 8 at :227).
static int block_cgrq_run(Op *op, Field *B, Field *X, double tol, int maxit, gx_solver_stats *st) {
  GX_TRY(op_prepare(op, GX_F64));
  Field *R = op_tmp(op, GX_F64, 4), *Q = op_tmp(op, GX_F64, 5), *D = op_tmp(op, GX_F64, 6), *Z = op_tmp(op, GX_F64, 9);
  if (!R || !Q || !D || !Z) return GX_ENOMEM;
  X->cb = R->cb = Q->cb = D->cb = Z->cb = B->cb;
  BlockWork bw;
  GX_TRY(bw.init(op->ctx));
  SliceBlock blk{op, &bw};
  blk.v[BV_B] = B, blk.v[BV_X] = X, blk.v[BV_R] = R, blk.v[BV_Q] = Q, blk.v[BV_D] = D, blk.v[BV_Z] = Z;
  return block_cgrq_core(blk, tol, maxit, st);
}

// BlockCGrQsolveVec over n FP64 fields of one subset: half fields with A = MpcDagMpc, or full
// fields with A = Mdag M.  The work fields (R, Q, D, Z: 4 n fields) live for the solve.
static int block_cgrq_vec_run(Op *op, int herm, int n, Field **B, Field **X, double tol, int maxit,
                              gx_solver_stats *st) {
  GX_TRY(op_prepare(op, GX_F64));
  BlockWork bw;
  GX_TRY(bw.init(op->ctx));
  VecBlock blk{op, &bw, herm, n};
  struct Owned {
    std::vector<Field *> list;
    ~Owned() {
      for (Field *f : list) field_destroy(f);
    }
  } owned;
  const int subset = herm == GX_HERM_SCHUR ? GX_EVEN : GX_FULL;
  for (int b = 0; b < n; b++) {
    X[b]->cb = B[b]->cb;
    blk.v[BV_B].push_back(B[b]);
    blk.v[BV_X].push_back(X[b]);
    for (int id = BV_R; id < BV_COUNT; id++) {
      Field *f = nullptr;
      GX_TRY(field_create(op->ctx, op->g, GX_F64, subset, &f));
      owned.list.push_back(f);
      f->cb = B[b]->cb;
      blk.v[id].push_back(f);
    }
  }
  if (herm == GX_HERM_MDAGM) {
    GX_TRY(field_create(op->ctx, op->g, GX_F64, GX_FULL, &blk.full_tmp));
    owned.list.push_back(blk.full_tmp);
  }
  const int rc = block_cgrq_core(blk, tol, maxit, st);
  GX_CHECK_HIP(hipStreamSynchronize(op->ctx->stream));  // the work fields are freed on return
  return rc;
}

// ConjugateGradientMultiShift (ConjugateGradientMultiShift.h:89-343): (MpcDagMpc + m_s) psi_s = src
// for all shifts m_s >= m_0 from the Krylov space of the lightest (Jegerlehner's shifted CG).  The
// scalar recurrences (a, b, z_s, bs_s) run in FP64 in Grid's order: the start on the host, every
// iteration after it in device finalisers (MsState), so the host never waits inside an iteration;
// psi_s start at zero; a shift stops updating once c z_s^2 < m_residual_s^2 |src|^2.  FP64 fields.
static int multishift_run(Op *op, Field *src, int nshift, const double *mass, const double *mres, Field **psi,
                          int maxit, gx_solver_stats *st) {
  const auto t0 = std::chrono::steady_clock::now();
  GX_TRY(op_prepare(op, GX_F64));
  Field *r = op_tmp(op, GX_F64, 4), *p = op_tmp(op, GX_F64, 5), *mmp = op_tmp(op, GX_F64, 6),
        *tmp = op_tmp(op, GX_F64, 9);
  if (!r || !p || !mmp || !tmp) return GX_ENOMEM;
  const int cb = src->cb;
  r->cb = p->cb = mmp->cb = tmp->cb = cb;
  std::vector<Field *> ps(nshift, nullptr);
  struct Owned {
    std::vector<Field *> &v;
    ~Owned() {
      for (Field *f : v)
        if (f) field_destroy(f);
    }
  } owned{ps};
  for (int s = 0; s < nshift; s++) {
    GX_TRY(field_create(op->ctx, op->g, GX_F64, GX_EVEN, &ps[s]));
    ps[s]->cb = cb;
    psi[s]->cb = cb;
  }
  hipStream_t stream = op->ctx->stream;
  std::vector<double> bs(nshift), rsq(nshift), zc(nshift), zp(nshift);  // z[s][iz], z[s][1-iz]
  std::vector<int> conv(nshift, 0), its(nshift, 0);
  double cp;
  GX_TRY(blas_norm2(src, &cp));
  if (std::isnan(cp)) {
    set_error("multishift CG: NaN in source");
    return GX_ENAN;
  }
  if (cp == 0.0) {
    for (int s = 0; s < nshift; s++) {
      GX_CHECK_HIP(hipMemsetAsync(psi[s]->data, 0, psi[s]->bytes, stream));
      if (st) {
        st[s] = gx_solver_stats{};
        st[s].iterations = 1;
        st[s].converged = 1;
      }
    }
    GX_CHECK_HIP(hipStreamSynchronize(stream));
    return GX_OK;
  }
  for (int s = 0; s < nshift; s++) {
    rsq[s] = cp * mres[s] * mres[s];
    GX_CHECK_HIP(hipMemcpyAsync(ps[s]->data, src->data, src->bytes, hipMemcpyDeviceToDevice, stream));
  }
  GX_CHECK_HIP(hipMemcpyAsync(r->data, src->data, src->bytes, hipMemcpyDeviceToDevice, stream));
  GX_CHECK_HIP(hipMemcpyAsync(p->data, src->data, src->bytes, hipMemcpyDeviceToDevice, stream));
  // d = Re<p, (A + m_0) p>, b = -cp / d
  double d, rn, c;
  auto shifted_op = [&](double &dd) -> int {
    GX_TRY(apply_mpcdagmpc(op, p, mmp));
    GX_TRY(blas_inner_norm(p, mmp, &dd, &rn));  // Re<p|A p> and |p|^2
    GX_TRY(blas_axpby(mmp, mass[0], p, 1.0, mmp));
    dd += rn * mass[0];
    return GX_OK;
  };
  GX_TRY(shifted_op(d));
  double b = -cp / d;
  zp[0] = zc[0] = 1.0;
  bs[0] = b;
  for (int s = 1; s < nshift; s++) {
    zp[s] = 1.0;
    zc[s] = 1.0 / (1.0 - b * (mass[s] - mass[0]));
    bs[s] = b * zc[s];
  }
  GX_TRY(blas_axpy_norm(r, b, mmp, r, &c));
  for (int s = 0; s < nshift; s++) GX_TRY(blas_axpby(psi[s], -bs[s], src, 0.0, src));
  // The vector updates of every shift run in one pass over r (k_multishift): the one that closes
  // iteration k (psi_s -= bs_s ps_s) is fused with the one that opens k + 1 (p = a p + r,
  // ps_s = z_s r + as_s ps_s), with Grid's coefficients.
  MultiShiftArgs m{};
  m.nshift = nshift;
  for (int s = 0; s < nshift; s++) {
    m.ps[s] = ps[s]->data;
    m.psi[s] = psi[s]->data;
  }
  const int64_t nst = src->stored();
  auto update = [&](uint32_t close, uint32_t open, double an) -> int {
    m.do_psi = close;
    m.do_ps = open;
    m.do_p = open != 0;
    m.a = an;
    for (int s = 0; s < nshift; s++) {
      m.bs[s] = bs[s];
      m.z[s] = s == 0 ? 1.0 : zc[s];
      m.as[s] = s == 0 ? an : an * zc[s] * bs[s] / (zp[s] * b);
    }
    return launch_multishift(op->ctx, nst, r->data, p->data, m);
  };
  uint32_t active = nshift == 32 ? 0xffffffffu : ((1u << nshift) - 1u);
  double a = c / cp;  // the a of the iteration being opened
  GX_TRY(update(0u, active, a));  // iteration 1: p and ps_s
  // Iterations 1, 2, ... with the recurrences on the device (k_ms_fin_d, k_ms_fin_c) and batches
  // enqueued two ahead, as in CG: the host only reads the state between batches.
  SolverWork *sw = nullptr;
  GX_TRY(solver_work(op, &sw));
  MsWork &w = sw->ms;
  if (!w.dev) {
    hipError_t e = hipMalloc(&w.dev, sizeof(MsState));
    if (e == hipSuccess) e = hipHostMalloc(&w.host, 2 * sizeof(MsState), hipHostMallocDefault);
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = hipEventCreateWithFlags(&w.ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
      set_error("multishift state: %s", hipGetErrorString(e));
      return GX_ENOMEM;
    }
  }
  MsState init{};
  init.k = 1;
  init.k_conv = INT_MAX;
  init.kmax = maxit;
  init.c = c;
  init.cp = cp;
  init.a = a;
  init.b = b;
  init.m0 = mass[0];
  for (int s = 0; s < nshift; s++) {
    init.zc[s] = zc[s];
    init.zp[s] = zp[s];
    init.bs[s] = bs[s];
    init.dm[s] = mass[s] - mass[0];
    init.rsq[s] = rsq[s];
  }
  init.active = active;
  init.closing = 0;
  init.nshift = nshift;
  w.host[0] = init;
  GX_CHECK_HIP(hipMemcpyAsync(w.dev, &w.host[0], sizeof(MsState), hipMemcpyHostToDevice, stream));
  GX_CHECK_HIP(hipStreamSynchronize(stream));
  MsPtrs ptrs{};
  for (int s = 0; s < nshift; s++) {
    ptrs.ps[s] = ps[s]->data;
    ptrs.psi[s] = psi[s]->data;
  }
  Ctx *ctx = op->ctx;
  const int64_t nsites = r->g.half_pad();
  const int64_t vol = op->g.half_elems();
  const int batch = vol >= (1 << 20) ? 8 : (vol >= (1 << 16) ? 16 : 32);
  int enqueued = 0, slot = 0;
  auto enqueue = [&](int sl) -> int {
    for (int i = 0; i < batch && enqueued < maxit; i++, enqueued++) {
      GX_TRY(apply_mpcdagmpc(op, p, mmp, w.dev));
      GX_TRY(launch_ms_dot(ctx, nsites, nsites, p->data, mmp->data, w.dev));
      GX_TRY(blas_axpby(mmp, mass[0], p, 1.0, mmp, w.dev));
      GX_TRY(launch_ms_axpy_norm(ctx, nsites, nsites, r->data, mmp->data, w.dev));
      GX_TRY(launch_ms_update(ctx, nst, r->data, p->data, ptrs, w.dev));
    }
    GX_CHECK_HIP(hipMemcpyAsync(&w.host[sl], w.dev, sizeof(MsState), hipMemcpyDeviceToHost, stream));
    GX_CHECK_HIP(hipEventRecord(w.ev[sl], stream));
    return GX_OK;
  };
  GX_TRY(enqueue(0));
  GX_TRY(enqueue(1));
  MsState fin{};
  for (;;) {
    GX_CHECK_HIP(hipEventSynchronize(w.ev[slot]));
    fin = w.host[slot];
    if (fin.k_conv != INT_MAX || fin.k > maxit) break;
    if (enqueued >= maxit) {
      GX_CHECK_HIP(hipEventSynchronize(w.ev[1 - slot]));
      fin = w.host[1 - slot];
      break;
    }
    GX_TRY(enqueue(slot));
    slot = 1 - slot;
  }
  // the final state (launches past the stop leave it as the stop found it, but for k)
  GX_CHECK_HIP(hipMemcpyAsync(&w.host[0], w.dev, sizeof(MsState), hipMemcpyDeviceToHost, stream));
  GX_CHECK_HIP(hipStreamSynchronize(stream));
  fin = w.host[0];
  int rc = fin.active == 0 ? GX_OK : GX_ENOCONV;
  for (int s = 0; s < nshift; s++) {
    its[s] = fin.its[s];
    conv[s] = ((fin.active >> s) & 1u) ? 0 : 1;
    zc[s] = fin.zc[s];
  }
  c = fin.c;
  // true residuals |(A + m_s) psi_s - src| / |src|
  double ssq;
  GX_TRY(blas_norm2(src, &ssq));
  const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  for (int s = 0; s < nshift; s++) {
    double rr;
    GX_TRY(apply_mpcdagmpc(op, psi[s], mmp));
    GX_TRY(blas_axpby(tmp, mass[s], psi[s], 1.0, mmp));
    GX_TRY(blas_axpy_norm(r, -1.0, src, tmp, &rr));
    if (st) {
      st[s] = gx_solver_stats{};
      st[s].iterations = its[s];
      st[s].converged = conv[s];
      st[s].true_residual = std::sqrt(rr / ssq);
      st[s].computed_residual = std::sqrt(std::max(0.0, c * zc[s] * zc[s]) / ssq);
      st[s].seconds = secs;
    }
  }
  return rc;
}

}  // namespace gx

using namespace gx;


extern "C" {

int gx_cg(gx_op op, gx_field src, gx_field psi, double tol, int maxit, gx_solver_stats *st) {
  if (!op || !src || !psi || src == psi) {
    set_error("cg: bad arguments");
    return GX_EINVAL;
  }
  Field *s = src, *p = psi;
  if (s->subset == GX_FULL || p->subset == GX_FULL || s->prec != p->prec || !same_geom(s->g, op->g) ||
      !same_geom(p->g, op->g)) {
    set_error("cg: needs half fields of one precision on the operator's lattice");
    return GX_EINVAL;
  }
  return cg_run(op, s, p, tol, maxit, st);
}

int gx_mixed_cg(gx_op op, gx_field src, gx_field psi, double tol, double inner_tol, int inner_prec, int maxinner,
                int maxouter, gx_solver_stats *st) {
  if (!op || !src || !psi || src == psi || src->prec != GX_F64 || psi->prec != GX_F64 ||
      src->subset == GX_FULL || psi->subset == GX_FULL || !same_geom(src->g, op->g) || !same_geom(psi->g, op->g)) {
    set_error("mixed_cg: needs F64 half fields on the operator's lattice");
    return GX_EINVAL;
  }
  if (inner_prec != GX_F32 && inner_prec != GX_F16) {
    set_error("mixed_cg: inner precision %d unsupported", inner_prec);
    return GX_EINVAL;
  }
  return mixed_cg_run(op, src, psi, tol, inner_tol, inner_prec, maxinner, maxouter, st);
}

int gx_reliable_update_cg(gx_op op, gx_field src, gx_field psi, double tol, int maxit, double delta, int sloppy_prec,
                          gx_solver_stats *st) {
  if (!op || !src || !psi || src == psi || src->prec != GX_F64 || psi->prec != GX_F64 || src->subset == GX_FULL ||
      psi->subset == GX_FULL || !same_geom(src->g, op->g) || !same_geom(psi->g, op->g)) {
    set_error("reliable_update_cg: needs F64 half fields on the operator's lattice");
    return GX_EINVAL;
  }
  if (!(delta > 0.0 && delta < 1.0) || (sloppy_prec != GX_F32 && sloppy_prec != GX_F16)) {
    set_error("reliable_update_cg: needs 0 < delta < 1 and a GX_F32 or GX_F16 sloppy precision");
    return GX_EINVAL;
  }
  return ru_cg_run(op, src, psi, tol, maxit, delta, sloppy_prec, -1, 0.0, st);
}

int gx_reliable_update_cg_fallback(gx_op op, gx_field src, gx_field psi, double tol, int maxit, double delta,
                                   int sloppy_prec, int fallback_prec, double fallback_transition_tol,
                                   gx_solver_stats *st) {
  if (sloppy_prec != GX_F16 || fallback_prec != GX_F32 || !(fallback_transition_tol >= 0.0)) {
    set_error("reliable_update_cg_fallback: needs GX_F16 sloppy, GX_F32 fallback precision and a transition "
              "tolerance >= 0");
    return GX_EINVAL;
  }
  if (!op || !src || !psi || src == psi || src->prec != GX_F64 || psi->prec != GX_F64 || src->subset == GX_FULL ||
      psi->subset == GX_FULL || !same_geom(src->g, op->g) || !same_geom(psi->g, op->g)) {
    set_error("reliable_update_cg_fallback: needs F64 half fields on the operator's lattice");
    return GX_EINVAL;
  }
  if (!(delta > 0.0 && delta < 1.0)) {
    set_error("reliable_update_cg_fallback: needs 0 < delta < 1");
    return GX_EINVAL;
  }
  return ru_cg_run(op, src, psi, tol, maxit, delta, sloppy_prec, fallback_prec, fallback_transition_tol, st);
}

int gx_block_cg(gx_op op_, gx_field src_, gx_field psi_, double tol, int maxit, int inner_prec,
                gx_solver_stats *st) {
  Op *op = op_;
  Field *src = src_, *psi = psi_;
  if (!op || !src || !psi || src == psi || op->kind != GX_WILSON || src->subset == GX_FULL ||
      psi->subset == GX_FULL || src->prec != GX_F64 || psi->prec != GX_F64 || !same_geom(src->g, op->g) ||
      !same_geom(psi->g, op->g)) {
    set_error("block_cg: needs a multi-RHS Wilson operator and F64 half fields");
    return GX_EINVAL;
  }
  if (inner_prec == GX_F64) return mrhs_cg_run(op, src, psi, tol, maxit, st);
  if (inner_prec == GX_F32 || inner_prec == GX_F16)
    return mrhs_mixed_run(op, src, psi, tol, inner_prec == GX_F32 ? 1e-6 : 1e-3, inner_prec, maxit, 100, st, -1,
                          kConvergeRule);
  set_error("block_cg: inner precision %d unsupported", inner_prec);
  return GX_EINVAL;
}

int gx_cg_multishift(gx_op op_, gx_field src_, int nshift, const double *shifts, const double *tolerances,
                     gx_field *psi_, int maxit, gx_solver_stats *st) {
  Op *op = op_;
  Field *src = src_;
  if (!op || !src || src->prec != GX_F64 || src->subset == GX_FULL || !same_geom(src->g, op->g) || nshift < 1 ||
      nshift > kMaxShift || !shifts || !tolerances || !psi_) {
    set_error("cg_multishift: needs an F64 half source on the operator's lattice and 1..%d shifts", kMaxShift);
    return GX_EINVAL;
  }
  std::vector<Field *> psi(nshift);
  for (int s = 0; s < nshift; s++) {
    psi[s] = psi_[s];
    if (!psi[s] || psi[s] == src || psi[s]->prec != GX_F64 || psi[s]->subset == GX_FULL ||
        !same_geom(psi[s]->g, op->g)) {
      set_error("cg_multishift: solution %d must be an F64 half field on the operator's lattice", s);
      return GX_EINVAL;
    }
    for (int t = 0; t < s; t++)
      if (psi[t] == psi[s]) {
        set_error("cg_multishift: solutions %d and %d are the same field", t, s);
        return GX_EINVAL;
      }
    if (!(shifts[s] >= shifts[0]) || !(tolerances[s] > 0.0)) {
      set_error("cg_multishift: shift %d must be >= shift 0 (the lightest) with a positive tolerance", s);
      return GX_EINVAL;
    }
  }
  return multishift_run(op, src, nshift, shifts, tolerances, psi.data(), maxit, st);
}

int gx_block_cgrq(gx_op op_, gx_field src_, gx_field psi_, double tol, int maxit, gx_solver_stats *st) {
  Op *op = op_;
  Field *src = src_, *psi = psi_;
  if (!op || !src || !psi || src == psi || op->kind != GX_WILSON || src->subset == GX_FULL ||
      psi->subset == GX_FULL || src->prec != GX_F64 || psi->prec != GX_F64 || !same_geom(src->g, op->g) ||
      !same_geom(psi->g, op->g)) {
    set_error("block_cgrq: needs a multi-RHS Wilson operator and F64 half fields");
    return GX_EINVAL;
  }
  return block_cgrq_run(op, src, psi, tol, maxit, st);
}

int gx_mixed_cg_batched(gx_op op_, int n, gx_field *src_, gx_field *sol_, double tol, double inner_tol, int inner_prec,
                        int maxinner, int maxouter, int maxpatch, gx_solver_stats *st) {
  Op *op = op_;
  if (!op || !src_ || !sol_ || n < 1 || (inner_prec != GX_F32 && inner_prec != GX_F16)) {
    set_error("mixed_cg_batched: needs an operator, n >= 1 fields and a GX_F32 or GX_F16 inner precision");
    return GX_EINVAL;
  }
  std::vector<Field *> S(n), X(n);
  for (int k = 0; k < n; k++) {
    S[k] = src_[k];
    X[k] = sol_[k];
    for (Field *f : {S[k], X[k]})
      if (!f || f->prec != GX_F64 || f->subset == GX_FULL || !same_geom(f->g, op->g)) {
        set_error("mixed_cg_batched: field %d must be an F64 half field on the operator's lattice", k);
        return GX_EINVAL;
      }
    if (S[k]->cb != S[0]->cb) {
      set_error("mixed_cg_batched: sources on different checkerboards");
      return GX_ECB;
    }
  }
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      if (X[k] == S[j] || (j != k && X[k] == X[j])) {
        set_error("mixed_cg_batched: solutions must be distinct fields, apart from the sources");
        return GX_EINVAL;
      }
  return mixed_batched_run(op, n, S.data(), X.data(), tol, inner_tol, inner_prec, maxinner, maxouter, maxpatch, st);
}

int gx_block_cgrq_vec(gx_op op_, int herm, int n, gx_field *src_, gx_field *psi_, double tol, int maxit,
                      gx_solver_stats *st) {
  Op *op = op_;
  if (!op || !src_ || !psi_ || n < 1 || n > kMaxLs || (herm != GX_HERM_SCHUR && herm != GX_HERM_MDAGM)) {
    set_error("block_cgrq_vec: needs an operator, 1..%d fields and herm GX_HERM_SCHUR or GX_HERM_MDAGM", kMaxLs);
    return GX_EINVAL;
  }
  std::vector<Field *> B(n), X(n);
  const bool half = herm == GX_HERM_SCHUR;
  for (int b = 0; b < n; b++) {
    B[b] = src_[b];
    X[b] = psi_[b];
    for (Field *f : {B[b], X[b]})
      if (!f || f->prec != GX_F64 || (f->subset == GX_FULL) == half || !same_geom(f->g, op->g)) {
        set_error("block_cgrq_vec: field %d must be an F64 %s field on the operator's lattice", b,
                  half ? "half" : "full");
        return GX_EINVAL;
      }
    if (half && B[b]->cb != B[0]->cb) {
      set_error("block_cgrq_vec: sources on different checkerboards");
      return GX_ECB;
    }
  }
  for (int b = 0; b < n; b++)
    for (int c = 0; c < n; c++)
      if (X[b] == B[c] || (c != b && (X[b] == X[c]))) {
        set_error("block_cgrq_vec: solutions must be distinct fields, apart from the sources");
        return GX_EINVAL;
      }
  return block_cgrq_vec_run(op, herm, n, B.data(), X.data(), tol, maxit, st);
}

// ---------------- allocation-free slice inner-product matrix ----------------
// Workspace of gx_slice_inner_product_matrix_noalloc: the workgroup partials (the same partition as
// BlockWork's, so the sums match gx_slice_inner_product_matrix bit for bit) and the result matrix,
// in a fixed device array.  Each GPU the library runs on holds its own copy until the process
// ends; its address is looked up once per device with that device current.
constexpr int kScratchParts = kIpmParts;
constexpr size_t kScratchMat = (size_t)kMaxLs * kMaxLs * sizeof(double2);
constexpr size_t kScratchBytes = (size_t)(kScratchParts + 1) * kScratchMat;  // 8.4 MB
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void *> g_scratch_addr;  // per device, null until looked up

static int scratch_address(int device, void **out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (device < 0) {
    set_error("scratch: bad device %d", device);
    return GX_EINVAL;
  }
  if ((int)g_scratch_addr.size() <= device) g_scratch_addr.resize(device + 1, nullptr);
  if (!g_scratch_addr[device]) {
    int cur = 0;
    GX_CHECK_HIP(hipGetDevice(&cur));
    if (cur != device) GX_CHECK_HIP(hipSetDevice(device));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) (void)hipSetDevice(cur);
    if (e != hipSuccess || !p) {
      set_error("scratch: hipGetSymbolAddress failed on device %d: %s", device, hipGetErrorString(e));
      return GX_EHIP;
    }
    g_scratch_addr[device] = p;
  }
  *out = g_scratch_addr[device];
  return GX_OK;
}

static bool slice_conformable(const Field *a, const Field *b) {
  return a && b && a->prec == GX_F64 && b->prec == GX_F64 && same_geom(a->g, b->g) &&
         (a->subset == GX_FULL) == (b->subset == GX_FULL);
}

int gx_slice_inner_product_matrix(gx_field lhs, gx_field rhs, double *mat) {
  if (!slice_conformable(lhs, rhs) || !mat) {
    set_error("slice_inner_product_matrix: needs two F64 fields of one lattice and subset");
    return GX_EINVAL;
  }
  BlockWork bw;
  GX_TRY(bw.init(lhs->ctx));
  CMat m;
  GX_TRY(bw.ipm(lhs->g, lhs->subset == GX_FULL ? 2 : 1, lhs, rhs, m));
  for (int i = 0; i < m.n * m.n; i++) {
    mat[2 * i] = m.a[i].real();
    mat[2 * i + 1] = m.a[i].imag();
  }
  return GX_OK;
}

int gx_slice_inner_product_matrix_noalloc(gx_field lhs, gx_field rhs, double *mat) {
  if (!slice_conformable(lhs, rhs) || !mat || lhs->g.Ls > kMaxLs) {
    set_error("slice_inner_product_matrix_noalloc: needs two F64 fields of one lattice and subset, Ls <= %d", kMaxLs);
    return GX_EINVAL;
  }
  Ctx *ctx = lhs->ctx;
  void *base = nullptr;
  GX_TRY(scratch_address(ctx->device, &base));
  void *partial = base, *out_dev = (char *)base + (size_t)kScratchParts * kScratchMat;
  const int n = lhs->g.Ls;
  GX_TRY(launch_slice_ipm(ctx, lhs->g, lhs->subset == GX_FULL ? 2 : 1, lhs->data, rhs->data, partial, kScratchParts,
                          out_dev));
  GX_CHECK_HIP(hipMemcpyAsync(mat, out_dev, sizeof(double2) * n * n, hipMemcpyDeviceToHost, ctx->stream));
  GX_CHECK_HIP(hipStreamSynchronize(ctx->stream));
  return GX_OK;
}

int gx_slice_madd_matrix(gx_field R, const double *aa, gx_field X, gx_field Y, double scale) {
  if (!slice_conformable(R, X) || (Y && !slice_conformable(R, Y)) || !aa) {
    set_error("slice_madd_matrix: needs F64 fields of one lattice and subset");
    return GX_EINVAL;
  }
  const int n = X->g.Ls;
  CMat a(n);
  for (int i = 0; i < n * n; i++) a.a[i] = cd(aa[2 * i], aa[2 * i + 1]);
  BlockWork bw;
  GX_TRY(bw.init(X->ctx));
  GX_TRY(bw.madd(X->g, X->subset == GX_FULL ? 2 : 1, R, a, X, Y, scale, 1));
  GX_CHECK_HIP(hipStreamSynchronize(X->ctx->stream));  // the workspace is freed on return
  return GX_OK;
}

// n fields of one lattice, subset and precision (F64) as a pointer list; false if they are not
static bool vec_fields(int n, gx_field *in, VecPtrs &p, const Field **first) {
  if (n < 1 || n > kMaxLs || !in) return false;
  p = VecPtrs{};
  for (int b = 0; b < n; b++) {
    const Field *f = in[b];
    if (!f || f->prec != GX_F64) return false;
    if (b > 0 && (!same_geom(f->g, in[0]->g) || (f->subset == GX_FULL) != (in[0]->subset == GX_FULL))) return false;
    p.p[b] = f->data;
  }
  *first = in[0];
  return true;
}

int gx_inner_product_matrix(int n, gx_field *X, gx_field *Y, double *mat) {
  VecPtrs px, py;
  const Field *fx = nullptr, *fy = nullptr;
  if (!mat || !vec_fields(n, X, px, &fx) || !vec_fields(n, Y, py, &fy) || !same_geom(fx->g, fy->g) ||
      (fx->subset == GX_FULL) != (fy->subset == GX_FULL)) {
    set_error("inner_product_matrix: needs 1..%d F64 fields of one lattice and subset on each side", kMaxLs);
    return GX_EINVAL;
  }
  BlockWork bw;
  GX_TRY(bw.init(fx->ctx));
  CMat m;
  GX_TRY(bw.ipm_vec(n, fx->stored(), px, py, m));
  for (int i = 0; i < n * n; i++) {
    mat[2 * i] = m.a[i].real();
    mat[2 * i + 1] = m.a[i].imag();
  }
  return GX_OK;
}

int gx_madd_matrix(int n, gx_field *R, const double *aa, gx_field *X, gx_field *Y, double scale) {
  VecPtrs pr, px, py{};
  const Field *fr = nullptr, *fx = nullptr, *fy = nullptr;
  bool ok = aa && vec_fields(n, R, pr, &fr) && vec_fields(n, X, px, &fx) && same_geom(fr->g, fx->g) &&
            (fr->subset == GX_FULL) == (fx->subset == GX_FULL);
  if (ok && Y) ok = vec_fields(n, Y, py, &fy) && same_geom(fr->g, fy->g) && (fr->subset == GX_FULL) == (fy->subset == GX_FULL);
  if (!ok) {
    set_error("madd_matrix: needs 1..%d F64 fields of one lattice and subset in R, X (and Y)", kMaxLs);
    return GX_EINVAL;
  }
  CMat a(n);
  for (int i = 0; i < n * n; i++) a.a[i] = cd(aa[2 * i], aa[2 * i + 1]);
  BlockWork bw;
  GX_TRY(bw.init(fx->ctx));
  GX_TRY(bw.madd_vec(n, fx->stored(), pr, a, px, py, scale, 1));
  for (int b = 0; b < n; b++) R[b]->cb = X[b]->cb;
  GX_CHECK_HIP(hipStreamSynchronize(fx->ctx->stream));  // the workspace is freed on return
  return GX_OK;
}

int gx_schur_solve(gx_op op_, gx_field src_, gx_field sol_, double tol, int maxit, int mixed, gx_solver_stats *st,
                   double *unprec_residual) {
  Op *op = op_;
  Field *src = src_, *sol = sol_;
  if (!op || !src || !sol || src == sol || src->subset != GX_FULL || sol->subset != GX_FULL || src->prec != GX_F64 ||
      sol->prec != GX_F64 || !same_geom(src->g, op->g) || !same_geom(sol->g, op->g)) {
    set_error("schur_solve: needs two distinct full F64 fields on the operator's lattice");
    return GX_EINVAL;
  }
  // the operator's own temporaries: repeated solves allocate (and free) nothing
  Field *se = op_tmp(op, GX_F64, 10), *so = op_tmp(op, GX_F64, 11), *t = op_tmp(op, GX_F64, 12);
  Field *Mt = op_tmp(op, GX_F64, 13), *solo = op_tmp(op, GX_F64, 14);
  if (!se || !so || !t || !Mt || !solo) return GX_ENOMEM;
  const bool multi = op->kind == GX_WILSON && op->g.Ls > 1;
  const int Ls = op->g.Ls;
  if (mixed < 0 || mixed > 3 || (multi && mixed == 3)) {
    set_error("schur_solve: mixed must be 0 (FP64), 1 (FP32 inner), 2 (FP16 inner) or, on single-RHS operators, "
              "3 (reliable updates)");
    return GX_EINVAL;
  }
  hipStream_t s = op->ctx->stream;
  const size_t hb = se->bytes;
  int rc = GX_OK;
  do {
    // RedBlackSource: src_o = MpcDag(src_o - Moe Mee^-1 src_e)
    if ((rc = (hipMemcpyAsync(se->data, src->half(0), hb, hipMemcpyDeviceToDevice, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    if ((rc = (hipMemcpyAsync(so->data, src->half(1), hb, hipMemcpyDeviceToDevice, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    se->cb = 0;
    so->cb = 1;
    if ((rc = apply_mooee(op, se, t, 2))) break;          // tmp = MooeeInv src_e   (Even)
    if ((rc = apply_meooe(op, t, Mt, 0))) break;          // Mtmp = Meooe tmp       (Odd)
    if ((rc = blas_axpby(t, 1.0, so, -1.0, Mt))) break;   // tmp = src_o - Mtmp     (Odd)
    if ((rc = apply_mpc(op, t, so, 1))) break;            // src_o = MpcDag tmp
    so->cb = 1;
    // RedBlackSolve with the zero guess
    solo->cb = 1;
    if ((rc = (hipMemsetAsync(solo->data, 0, solo->bytes, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    if (multi) {
      // Ls independent right-hand sides: every one to tol (CGmultiRHS).  First-leg inner tolerance:
      // 1e-6 with FP32 inner solves, so that the second leg's FP32 error stays below a 1e-11 stop
      // and two legs suffice (C4: 71 inner iterations in 2 legs against 73 in 3 with 1e-5, 9%
      // faster); 1e-3 with FP16 storage, whose legs gain three to four orders whatever they aim at
      // (C4: 72 inner iterations, 0.381 s, against 81, 0.409 s, at 1e-5).
      std::vector<gx_solver_stats> bst(Ls);
      rc = mixed ? mrhs_mixed_run(op, so, solo, tol, mixed == 2 ? 1e-3 : 1e-6, mixed == 2 ? GX_F16 : GX_F32, maxit,
                                  100, bst.data(), -1, kConvergeRule)
                 : mrhs_cg_run(op, so, solo, tol, maxit, bst.data());
      if (st) {
        *st = bst[0];
        for (int k = 1; k < Ls; k++) {
          st->iterations = std::max(st->iterations, bst[k].iterations);
          st->final_iterations = std::max(st->final_iterations, bst[k].final_iterations);
          st->converged = st->converged && bst[k].converged;
          st->true_residual = std::max(st->true_residual, bst[k].true_residual);
          st->computed_residual = std::max(st->computed_residual, bst[k].computed_residual);
        }
      }
    } else if (mixed == 3) {
      // ConjugateGradientReliableUpdate, FP32 sloppy vectors, delta 0.01: on the C3 system the
      // same 636 iterations with 15 FP64 updates instead of 31 at delta 0.1, 6% faster
      rc = ru_cg_run(op, so, solo, tol, maxit, 0.01, GX_F32, -1, 0.0, st);
    } else if (mixed) {
      // Grid's outer rule: on the ill-conditioned single-RHS Mobius systems the FP32 policy of the
      // multi-RHS solves took more inner iterations (C2 558 against 482, C3 694 against 594).
      // InnerTolerance 1e-5 with FP32 inner solves; 1e-3 with FP16 storage, whose legs gain about
      // three orders whatever they aim at (C3: 929 inner iterations in 9 legs, 0.77 s, against
      // 1117 in 7, 0.92 s, at 1e-5)
      rc = mixed_cg_run(op, so, solo, tol, mixed == 2 ? 1e-3 : 1e-5, mixed == 2 ? GX_F16 : GX_F32, maxit, 100, st);
    } else {
      rc = cg_run(op, so, solo, tol, maxit, st);
    }
    if (rc != GX_OK && rc != GX_ENOCONV) break;
    const int solve_rc = rc;
    // RedBlackSolution: sol_e = MooeeInv(src_e - Meo sol_o)
    if ((rc = apply_meooe(op, solo, t, 0))) break;        // tmp = Meooe sol_o (Even)
    if ((rc = (hipMemcpyAsync(se->data, src->half(0), hb, hipMemcpyDeviceToDevice, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    if ((rc = blas_axpby(Mt, 1.0, se, -1.0, t))) break;   // src_e - tmp
    Mt->cb = 0;
    if ((rc = apply_mooee(op, Mt, t, 2))) break;          // sol_e
    if ((rc = (hipMemcpyAsync(sol->half(0), t->data, hb, hipMemcpyDeviceToDevice, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    if ((rc = (hipMemcpyAsync(sol->half(1), solo->data, hb, hipMemcpyDeviceToDevice, s) == hipSuccess) ? GX_OK : GX_EHIP)) break;
    if (unprec_residual) {
      if (!op->full_tmp && (rc = field_create(op->ctx, op->g, GX_F64, GX_FULL, &op->full_tmp))) break;
      Field *res = op->full_tmp;
      if ((rc = apply_full_m(op, sol, res, 0))) break;
      if ((rc = blas_axpby(res, 1.0, res, -1.0, src))) break;
      if (multi) {
        // largest per-right-hand-side |M sol_k - src_k| / |src_k|, both checkerboards
        std::vector<double> r0(Ls), r1(Ls), s0(Ls), s1(Ls);
        Field rh0 = half_view(res, 0), rh1 = half_view(res, 1), sh0 = half_view(src, 0), sh1 = half_view(src, 1);
        if ((rc = slice_reduce(op, 1, &rh0, nullptr, nullptr, r0.data()))) break;
        if ((rc = slice_reduce(op, 1, &rh1, nullptr, nullptr, r1.data()))) break;
        if ((rc = slice_reduce(op, 1, &sh0, nullptr, nullptr, s0.data()))) break;
        if ((rc = slice_reduce(op, 1, &sh1, nullptr, nullptr, s1.data()))) break;
        double worst = 0;
        for (int k = 0; k < Ls; k++) {
          const double den = s0[k] + s1[k];
          worst = std::max(worst, den > 0 ? std::sqrt((r0[k] + r1[k]) / den) : std::sqrt(r0[k] + r1[k]));
        }
        *unprec_residual = worst;
      } else {
        double nr = 0, ns = 0;
        if ((rc = blas_norm2(res, &nr))) break;
        if ((rc = blas_norm2(src, &ns))) break;
        *unprec_residual = std::sqrt(nr / ns);
      }
    }
    rc = solve_rc;
  } while (0);
  if (hipStreamSynchronize(s) != hipSuccess && (rc == GX_OK || rc == GX_ENOCONV)) rc = GX_EHIP;
  return rc;
}

}  // extern "C"
