/*
 * gx.h — C ABI of the MI355X (gfx950) Wilson / Mobius domain-wall Dirac operator and
 * even-odd conjugate-gradient library (libgx.so).
 *
 * Every entry point returns an int status (GX_OK = 0); no C++ exception crosses the ABI.
 * All work is enqueued on the context's HIP stream.  Calls that hand back host scalars
 * (norms, inner products, solver statistics) synchronise that stream; the others do not.
 * gx_ctx_sync() waits for everything enqueued so far.
 *
 * Citations name the Grid (paboyle/Grid) interface each entry point replaces; paths are
 * relative to the Grid source tree.
 *
 * Host field layout (import/export): Grid's lexicographic site order, x fastest, with the
 * fifth dimension s innermost, and complex numbers as interleaved (re, im) doubles:
 *   full field : psi[((t*Lz + z)*Ly + y)*Lx + x][s][spin][colour]
 *   half field : psi[((t*Lz + z)*Ly + y)*(Lx/2) + x/2][s][spin][colour]   (one checkerboard)
 *   gauge      : U[site][mu][a][b], mu = 0..3 for x,y,z,t
 * Checkerboard (parity) of a site = (x+y+z+t) & 1; Even = 0, Odd = 1
 * (Grid/cartesian/Cartesian_red_black.h:39-95; 5D: Grid/qcd/utils/SpaceTimeGrid.cc:66-78).
 */
#ifndef GX_H
#define GX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum gx_status {
  GX_OK = 0,
  GX_EINVAL = 1,    /* bad argument / nonconformable fields (Grid: conformable() assert) */
  GX_ENOMEM = 2,    /* device allocation failed, or serving admission refused */
  GX_ENOCONV = 3,   /* solver did not converge (Grid: ErrorOnNoConverge assert) */
  GX_EHIP = 4,      /* HIP runtime error */
  GX_ECB = 5,       /* wrong checkerboard on input (Grid: assert(in.Checkerboard()==...)) */
  GX_ENAN = 6       /* NaN/Inf detected in an input (Grid: assert(std::isnan(guess)==0)) */
};

enum gx_prec { GX_F64 = 0, GX_F32 = 1, GX_F16 = 2 };   /* storage precision of a field */
enum gx_subset { GX_FULL = -1, GX_EVEN = 0, GX_ODD = 1 };
enum gx_kind { GX_WILSON = 0, GX_MOBIUS = 1 };

typedef struct gx_ctx_s *gx_ctx;
typedef struct gx_field_s *gx_field;
typedef struct gx_op_s *gx_op;

/* Last error message of the calling thread (static storage). */
const char *gx_last_error(void);
/* Library version string. */
const char *gx_version(void);

/* ---- context: device + stream (Grid: acceleratorInit, Grid/threads/Accelerator.cc:120-200).
 *      Destroy every field and operator created on a context before the context itself. ---- */
int gx_ctx_create(int device, gx_ctx *out);
int gx_ctx_destroy(gx_ctx ctx);
int gx_ctx_sync(gx_ctx ctx);
/* Timing on the context's stream: record an event pair and read the elapsed milliseconds. */
int gx_ctx_event_record(gx_ctx ctx, int slot);               /* slot 0..15 */
int gx_ctx_event_elapsed(gx_ctx ctx, int slot0, int slot1, double *ms);

/* ---- fields (Grid: Lattice<vobj> on a GridCartesian / GridRedBlackCartesian,
 *      Grid/lattice/Lattice_base.h:46; EO pick/set Grid/lattice/Lattice_transfer.h:50-80) ---- */
/* subset: GX_FULL for the whole lattice, GX_EVEN / GX_ODD for a half (checkerboarded) field. */
int gx_field_create(gx_ctx ctx, const int dims[4], int Ls, int prec, int subset, gx_field *out);
int gx_field_destroy(gx_field f);
int gx_field_import(gx_field f, const double *host);        /* host layout above */
int gx_field_export(gx_field f, double *host);
int gx_field_checkerboard(gx_field f, int *cb);              /* Grid: Lattice::Checkerboard() */
int gx_field_set_checkerboard_tag(gx_field f, int cb);       /* Grid: Checkerboard() = cb */
int gx_field_zero(gx_field f);                               /* Grid: field = Zero() */
int gx_field_copy(gx_field dst, gx_field src);               /* Grid: dst = src (same precision) */
/* Grid: pickCheckerboard(cb, half, full) / setCheckerboard(full, half), Lattice_transfer.h:50-80 */
int gx_pick_checkerboard(int cb, gx_field half, gx_field full);
int gx_set_checkerboard(gx_field full, gx_field half);
/* Grid: precisionChange(out, in), Lattice_transfer.h:1312-1492 */
int gx_precision_change(gx_field out, gx_field in);
/* Grid: ExtractSlice(lowDim, higherDim, slice, orthog) / InsertSlice(lowDim, higherDim, slice,
 * orthog), Lattice_transfer.h:949-1047, with orthog the fifth dimension: a 4D field (Ls = 1) from
 * or into slice s of a field on the same 4D lattice.  Two full or two half fields of one precision
 * (F64 or F32); the receiving half field takes the other's checkerboard. */
int gx_extract_slice(gx_field out4d, gx_field in5d, int s);
int gx_insert_slice(gx_field in4d, gx_field out5d, int s);

/* ---- BLAS1 / reductions (Grid/lattice/Lattice_arith.h:231-265; Lattice_reduction.h:202-410).
 *      Per-site products in the field's precision, global sums accumulated in double with a
 *      fixed-order two-stage reduction (bitwise reproducible run to run). ---- */
int gx_axpy(gx_field z, double a, gx_field x, gx_field y);                   /* z = a x + y */
int gx_axpby(gx_field z, double a, double b, gx_field x, gx_field y);        /* z = a x + b y */
int gx_axpy_norm(gx_field z, double a, gx_field x, gx_field y, double *nrm); /* z = a x + y, |z|^2 */
int gx_scale(gx_field z, double a, gx_field x);                              /* z = a x */
int gx_norm2(gx_field x, double *nrm);
int gx_inner_product(gx_field x, gx_field y, double *re, double *im);        /* <x|y> = sum conj(x) y */
int gx_inner_product_norm(gx_field x, gx_field y, double *re, double *im, double *nrm_x);

/* ---- fermion operators (Grid: WilsonFermion ctor WilsonFermion.h:139-142,
 *      MobiusFermion ctor MobiusFermion.h:45-51, ImportGauge FermionOperator.h:144).
 * U is the host gauge field (layout above); phases are the boundary phases as 4 (re,im)
 * pairs (NULL = periodic; Grid WilsonImplParams::boundary_phases, ActionParams.h:49-65).
 * For GX_WILSON, Ls > 1 gives Ls independent copies of the 4D operator sharing the gauge
 * links: the multi-right-hand-side operator (Grid: WilsonFermion5D with Ls = Nrhs,
 * WilsonFermion5D.h:40-65). ---- */
int gx_wilson_create(gx_ctx ctx, const int dims[4], int Ls, double mass, const double *U,
                     const double *phases, gx_op *out);
int gx_mobius_create(gx_ctx ctx, const int dims[4], int Ls, double mass, double M5, double b, double c,
                     const double *U, const double *phases, gx_op *out);
int gx_op_destroy(gx_op op);
int gx_op_import_gauge(gx_op op, const double *U);           /* Grid: ImportGauge */
int gx_op_set_mass(gx_op op, double mass);                   /* Grid: CayleyFermion5D::SetMass / new op */
/* No Grid counterpart (a Grid operator has one mass): a multi-right-hand-side Wilson operator
 * (Ls = Nrhs) with mass masses[s] on slice s, so solves of one gauge field at different masses
 * share one launch per hop (WilsonFermion::Mooee = 4 + m per slice, WilsonFermionImplementation.h:71).
 * n must equal Ls; masses = NULL returns to the operator's single mass.  gx_op_set_mass clears the
 * per-slice masses. */
int gx_op_set_slice_masses(gx_op op, int n, const double *masses);
int gx_op_slice_masses(gx_op op, int n, double *masses);     /* the mass of every slice (n = Ls) */
/* No Grid counterpart (Grid copies the gauge field into every operator, WilsonFermion.h:161-169):
 * a Wilson operator on base's 4D lattice that reads base's resident device links instead of
 * holding its own, with its own context (stream), Ls (right-hand sides) and mass, so concurrent
 * solves on one gauge field hold one copy of the links.  base must be a GX_WILSON operator
 * and must outlive the clone; gx_op_import_gauge on a clone is rejected. */
int gx_wilson_clone(gx_op base, gx_ctx ctx, int Ls, double mass, gx_op *out);
/* Re-point a clone at another base operator's links (same 4D lattice and boundary phases). */
int gx_op_share_links(gx_op clone, gx_op base);
/* Device bytes the operator holds now (links it owns, fifth-dimension matrices, temporaries). */
int gx_op_footprint(gx_op op, size_t *bytes);
/* No Grid counterpart (Grid's mixed-precision solvers trust the caller to pass the same action in
 * two precisions, ConjugateGradientMixedPrec.h:57-69): *same = 1 when a and b define the same
 * operator (kind, lattice, Ls, mass, M5, b, c, boundary phases and gauge field, bit for bit), so a
 * solver may run b's sloppy solves with a's links. */
int gx_op_same_action(gx_op a, gx_op b, int *same);

/* Operator applications; fields may be F64 or F32, in and out of the same precision.
 * Grid: FermionOperator.h:63-77; WilsonFermionImplementation.h:113-176,310-342;
 * CayleyFermion5DImplementation.h:155-329. */
int gx_M(gx_op op, gx_field in, gx_field out);
int gx_Mdag(gx_op op, gx_field in, gx_field out);
int gx_Meooe(gx_op op, gx_field in, gx_field out);
int gx_MeooeDag(gx_op op, gx_field in, gx_field out);
int gx_Mooee(gx_op op, gx_field in, gx_field out);
int gx_MooeeDag(gx_op op, gx_field in, gx_field out);
int gx_MooeeInv(gx_op op, gx_field in, gx_field out);
int gx_MooeeInvDag(gx_op op, gx_field in, gx_field out);
int gx_Dhop(gx_op op, gx_field in, gx_field out, int dag);     /* full lattice */
int gx_DhopEO(gx_op op, gx_field in, gx_field out, int dag);   /* odd in -> even out */
int gx_DhopOE(gx_op op, gx_field in, gx_field out, int dag);   /* even in -> odd out */
/* Grid: WilsonFermion::DhopDir(in, out, dir, disp) / Mdir (WilsonFermionImplementation.h:345-366;
 * WilsonFermion5D likewise): the one leg of the non-dagger hop in direction dir (0..3) with
 * displacement disp = +1 (out(x) = Uds_dir(x) (1 - gamma_dir) in(x + dir)) or -1 (out(x) =
 * Uds_{dir+4}(x) (1 + gamma_dir) in(x - dir)), Uds the -1/2-scaled double-stored links; the eight
 * legs sum to Dhop.  Full F64 fields. */
int gx_DhopDir(gx_op op, gx_field in, gx_field out, int dir, int disp);
/* Grid: DhopDeriv / DhopDerivEO / DhopDerivOE(mat, A, B, dag) (WilsonFermion5DImplementation.h:211-305;
 * WilsonImpl::InsertForce5D, WilsonImpl.h:193-238): for every mu, mat_mu(x) = sum over s and spin
 * of the colour outer product [Uds_mu(x) P B(x + mu)] A(x)^dag, P = (1 - gamma_mu) for dag = 0 and
 * (1 + gamma_mu) for dag = 1.  mat: host gauge layout (above); DhopDeriv fills every site from full
 * F64 fields A, B; DhopDerivEO (A even, B odd) the even sites and DhopDerivOE (A odd, B even) the
 * odd ones, leaving the others as they were. */
int gx_DhopDeriv(gx_op op, double *mat, gx_field A, gx_field B, int dag);
int gx_DhopDerivEO(gx_op op, double *mat, gx_field A, gx_field B, int dag);
int gx_DhopDerivOE(gx_op op, double *mat, gx_field A, gx_field B, int dag);
/* Cayley 5th-dimension pieces (Grid: CayleyFermion5D.h:43-116); Mobius operators only. */
int gx_M5D(gx_op op, gx_field psi, gx_field chi);       /* chi = M5D(psi, phi=chi) as Grid's M5D(psi,chi) */
int gx_M5Ddag(gx_op op, gx_field psi, gx_field chi);
int gx_Meooe5D(gx_op op, gx_field psi, gx_field chi);
int gx_MeooeDag5D(gx_op op, gx_field psi, gx_field chi);

/* 4D <-> 5D maps of a Mobius operator (full fields; the 4D field has Ls = 1 and the operator's
 * 4D dimensions).  Grid: CayleyFermion5D::ImportPhysicalFermionSource (P+ at s=0, P- at s=Ls-1,
 * then Dminus = 1 - c DW) and ExportPhysicalFermionSolution (P- psi(0) + P+ psi(Ls-1)),
 * CayleyFermion5DImplementation.h:57-129. */
int gx_import_physical_fermion_source(gx_op op, gx_field in4d, gx_field out5d);
int gx_export_physical_fermion_solution(gx_op op, gx_field sol5d, gx_field out4d);

/* ---- Schur even-odd operator (Grid: SchurDiagMooeeOperator, LinearOperator.h:325-349;
 *      SchurOperatorBase::MpcDagMpc/HermOpAndNorm :286-320) ---- */
int gx_Mpc(gx_op op, gx_field in, gx_field out);
int gx_MpcDag(gx_op op, gx_field in, gx_field out);
int gx_MpcDagMpc(gx_op op, gx_field in, gx_field out);
int gx_HermOpAndNorm(gx_op op, gx_field in, gx_field out, double *n1, double *n2);

/* ---- solvers ---- */
typedef struct {
  int iterations;          /* Grid: IterationsToComplete */
  int converged;
  double true_residual;    /* Grid: TrueResidual = |MpcDagMpc x - b| / |b| recomputed in FP64 */
  double computed_residual;/* sqrt(cp/ssq) at exit */
  double seconds;          /* wall time of the solve */
  int inner_iterations;    /* mixed precision: total inner iterations */
  int outer_iterations;    /* mixed precision: restarts; block CG: -- */
  int final_iterations;    /* mixed precision: FP64 patch-up iterations */
} gx_solver_stats;

/* Grid: ConjugateGradient<Field>(tol, maxit)(SchurDiagMooeeOperator(Dw), src, psi),
 * ConjugateGradient.h:43-257.  src, psi are half fields on the same checkerboard; psi holds
 * the initial guess. Returns GX_ENOCONV if not converged in maxit. */
int gx_cg(gx_op op, gx_field src, gx_field psi, double tol, int maxit, gx_solver_stats *st);

/* Grid: MixedPrecisionConjugateGradient<FieldD,FieldF>(tol, maxinner, maxouter, grid_f, Linop_f,
 * Linop_d)(src_d, sol_d), ConjugateGradientMixedPrec.h:57-165.  Inner solves run on fields of
 * precision inner_prec (GX_F32, or GX_F16 for FP16 storage with FP32 arithmetic) with
 * relative tolerance inner_tol; the outer residual, the solution accumulation and the final
 * patch-up CG are FP64.  src, psi: F64 half fields. */
int gx_mixed_cg(gx_op op, gx_field src, gx_field psi, double tol, double inner_tol, int inner_prec,
                int maxinner, int maxouter, gx_solver_stats *st);

/* Grid: ConjugateGradientReliableUpdate<FieldD,FieldF>(tol, maxit, delta, sp_grid, Linop_f,
 * Linop_d)(src, psi), ConjugateGradientReliableUpdate.h:77-269.  CG on MpcDagMpc whose Krylov
 * vectors live in sloppy_prec (GX_F32, or GX_F16 storage) while psi accumulates in FP64; when the
 * residual falls below delta times its largest value since the last update, the residual is
 * recomputed in FP64 from psi (a reliable update); a final FP64 CG from psi cleans up
 * (DoFinalCleanup).  st: inner_iterations = sloppy iterations, outer_iterations = reliable
 * updates, final_iterations = cleanup iterations.  src, psi: F64 half fields; 0 < delta < 1. */
int gx_reliable_update_cg(gx_op op, gx_field src, gx_field psi, double tol, int maxit, double delta,
                          int sloppy_prec, gx_solver_stats *st);

/* Grid: ConjugateGradientReliableUpdate::setFallbackLinop(Linop_fallback, fallback_transition_tol)
 * then operator(), ConjugateGradientReliableUpdate.h:72-75,255-259.  gx_reliable_update_cg with
 * GX_F16 sloppy vectors that switch to GX_F32 storage once the sloppy residual |r|^2 (absolute,
 * as in Grid) falls below fallback_transition_tol; p and r are promoted exactly and the iteration
 * continues.  FP16 Krylov vectors lose conjugacy on ill-conditioned systems and can stall far
 * above tol; the switch lets the FP16 phase take the cheap early iterations.  0 never switches. */
int gx_reliable_update_cg_fallback(gx_op op, gx_field src, gx_field psi, double tol, int maxit, double delta,
                                   int sloppy_prec, int fallback_prec, double fallback_transition_tol,
                                   gx_solver_stats *st);

/* Grid: SchurRedBlackDiagMooeeSolve<Field>(CG)(Dw, src, sol), SchurRedBlack.h:236-291,373-436.
 * src, sol: full F64 fields.  mixed = 1 (2) runs the red-black solve as gx_mixed_cg with FP32
 * (FP16-storage) inner solves from inner_tol 1e-5 (1e-3), mixed = 0 as FP64 gx_cg, mixed = 3
 * (single-RHS operators) as
 * gx_reliable_update_cg with FP32 sloppy vectors and delta 0.01.
 * unprec_residual = |M sol - src| / |src|.  On a multi-RHS Wilson operator (Ls > 1) the
 * red-black solve is gx_block_cg with inner_prec F64, F32 or F16 for mixed = 0, 1, 2 (every
 * right-hand side to tol), st holds the largest iteration
 * count and residuals over the right-hand sides, and unprec_residual is the largest
 * per-right-hand-side |M sol_k - src_k| / |src_k|. */
int gx_schur_solve(gx_op op, gx_field src, gx_field sol, double tol, int maxit, int mixed,
                   gx_solver_stats *st, double *unprec_residual);

/* Grid: BlockConjugateGradient(CGmultiRHS, Orthog = 0, tol, maxit)(Linop, src, psi),
 * BlockConjugateGradient.h:412-545: CG with per-right-hand-side alpha and beta on a multi-RHS
 * Wilson operator (GX_WILSON with Ls = Nrhs; right-hand side k is slice s = k), iterating every
 * right-hand side until the largest relative residual is below tol.  inner_prec = GX_F64 runs it
 * in FP64; GX_F32 (GX_F16) runs FP32 (FP16-storage) multi-RHS inner solves inside FP64 defect
 * correction (ConjugateGradientMixedPrecBatched.h:100-213, inner tolerance 1e-5) and an FP64
 * patch-up.
 * src, psi: F64 half fields; st: array of Ls statistics (one per right-hand side). */
int gx_block_cg(gx_op op, gx_field src, gx_field psi, double tol, int maxit, int inner_prec,
                gx_solver_stats *st);

/* Grid: BlockConjugateGradient(BlockCGrQ, Orthog = 0, tol, maxit)(Linop, src, psi),
 * BlockConjugateGradient.h:223-410: block CG on MpcDagMpc of a multi-RHS Wilson operator.  It
 * uses one block Krylov space for all Ls right-hand sides, with a thin QR of the residual block
 * (Cholesky of R^dag R) each iteration, and stops when the largest relative residual is below tol.
 * Nblock = Ls.  Grid's code fixes Nblock = 8 at :227, which is the same thing for Ls = 8.  Every
 * right-hand side must be nonzero, and the block must keep full rank.  src, psi: F64 half fields;
 * st: Ls statistics with per-right-hand-side true residuals.  FP64 only. */
int gx_block_cgrq(gx_op op, gx_field src, gx_field psi, double tol, int maxit, gx_solver_stats *st);

/* Grid: MixedPrecisionConjugateGradientBatched<FieldD,FieldF>(tol, maxinner, maxouter, maxpatchup,
 * sp_grid, Linop_f, Linop_d)(std::vector<FieldD> src, std::vector<FieldD> sol),
 * ConjugateGradientMixedPrecBatched.h:37-213: defect correction for n sources with FP64 outer
 * residuals, one inner tolerance for the batch (doubled while the largest residual times its
 * square is below the largest stopping norm, starting from inner_tol), inner solves in inner_prec
 * (GX_F32, or GX_F16 storage) and an FP64 patch-up CG of at most maxpatch iterations per source.
 * On a single-RHS Wilson operator and 2..16 sources the batch runs as one multi-RHS solve on a
 * clone of the operator (every link load shared by the batch); otherwise source by source.  src,
 * sol: n F64 half fields of one checkerboard; sol holds the guesses; st: n statistics. */
int gx_mixed_cg_batched(gx_op op, int n, gx_field *src, gx_field *sol, double tol, double inner_tol, int inner_prec,
                        int maxinner, int maxouter, int maxpatch, gx_solver_stats *st);

/* Grid: BlockConjugateGradient(BlockCGrQVec, _, tol, maxit)(Linop, std::vector<Field> src, psi),
 * BlockConjugateGradient.h:554-741 (BlockCGrQsolveVec): block CG over a list of n fields, with the
 * operator applied field by field.  herm = GX_HERM_SCHUR: A = MpcDagMpc (SchurDiagMooeeOperator) on
 * F64 half fields of one checkerboard; GX_HERM_MDAGM: A = Mdag M (MdagMLinearOperator) on full F64
 * fields.  Any operator kind (Wilson, Mobius).  psi holds the guesses.  Stops when the largest
 * relative residual of the block recursion is below tol; st: n statistics with true residuals
 * |A psi_b - src_b| / |src_b|.  1 <= n <= 16; every source must be nonzero. */
enum gx_herm { GX_HERM_SCHUR = 0, GX_HERM_MDAGM = 1 };
int gx_block_cgrq_vec(gx_op op, int herm, int n, gx_field *src, gx_field *psi, double tol, int maxit,
                      gx_solver_stats *st);

/* Grid: ConjugateGradientMultiShift<Field>(maxit, MultiShiftFunction)(Linop, src, results),
 * ConjugateGradientMultiShift.h:89-343: solves (MpcDagMpc + shifts[s]) psi[s] = src for nshift
 * shifts at once (poles = shifts, tolerances = per-shift relative residuals), from the Krylov
 * space of shifts[0], which must be the smallest.  Guesses are ignored (zero), as in Grid.
 * 1 <= nshift <= 32.  psi: nshift distinct F64 half fields; st: nshift statistics (iterations until each shift
 * converged, true residual |(A + shift) psi - src| / |src|).  GX_ENOCONV when a shift is not
 * converged after maxit iterations. */
int gx_cg_multishift(gx_op op, gx_field src, int nshift, const double *shifts, const double *tolerances,
                     gx_field *psi, int maxit, gx_solver_stats *st);

/* Grid: sliceInnerProductMatrix(mat, lhs, rhs, Orthog = 0), Lattice_reduction.h:757-776:
 * mat(s, t) = <lhs_s, rhs_t> over fifth-dimension slices s, t.  mat: 2 Ls^2 doubles (re, im),
 * row-major.  F64 fields of one lattice and subset. */
int gx_slice_inner_product_matrix(gx_field lhs, gx_field rhs, double *mat);

/* No counterpart in the reference (Grid allocates freely).  For callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that forbid hipMalloc):
 * gx_slice_inner_product_matrix with its workspace (workgroup partials and the result matrix,
 * 8.4 MB) in a fixed device array that each GPU holds until the process ends, instead of memory
 * allocated per call.  Same result, bit for bit.  Ls <= 16.  Calls that overlap on one GPU race on
 * that array. */
int gx_slice_inner_product_matrix_noalloc(gx_field lhs, gx_field rhs, double *mat);

/* Grid: sliceMaddMatrix(R, aa, X, Y, Orthog = 0, scale), Lattice_reduction.h:717-742:
 * R_t = Y_t + scale sum_s X_s aa(s, t).  aa: 2 Ls^2 doubles, row-major.  Y = NULL gives
 * sliceMulMatrix (R_t = scale sum_s X_s aa(s, t)).  R may be X or Y. */
int gx_slice_madd_matrix(gx_field R, const double *aa, gx_field X, gx_field Y, double scale);

/* Grid: InnerProductMatrix(m, X, Y) over std::vector<Field> (BlockConjugateGradient.h:35-42):
 * mat(b, bp) = <X_b, Y_bp> for n F64 fields of one lattice and subset.  mat: 2 n^2 doubles (re, im),
 * row-major.  1 <= n <= 16. */
int gx_inner_product_matrix(int n, gx_field *X, gx_field *Y, double *mat);

/* Grid: MaddMatrix(R, aa, X, Y, scale) / MulMatrix(R, aa, X) over std::vector<Field>
 * (BlockConjugateGradient.h:44-74): R_b = Y_b + scale sum_bp X_bp aa(bp, b); Y = NULL gives
 * MulMatrix.  R may be X or Y.  aa: 2 n^2 doubles, row-major. */
int gx_madd_matrix(int n, gx_field *R, const double *aa, gx_field *X, gx_field *Y, double scale);

/* ---- solve serving (no Grid counterpart) ----
 * A request asks for the Wilson solves M x_k = b_k, k < nsrc, on one of the resident gauge fields
 * of its lattice, with Gaussian sources b (gx_field_gaussian, Ls = nsrc, the request's seed).
 * gx_serve replays a trace of requests: request i becomes visible arrival[i] seconds after the
 * start (realtime = 1; all at the start with realtime = 0) and waits in FIFO order until a slot
 * is idle and its device-memory footprint fits the budget; it then runs as a red-black
 * (Schur) solve of all its sources at once (gx_schur_solve on a multi-RHS clone of the resident
 * operator, gx_wilson_clone) on that slot's own stream and host thread, overlapping on the device
 * with the solves of the other slots.  Resident gauge fields are built before the clock starts:
 * field g of a lattice comes from gx_random_gauge(seed = gauge_seed + g, eps = gauge_eps), or
 * from a NERSC file when gauge_dir is set. */
typedef struct {
  int dims[4];
  int nsrc;          /* right-hand sides, 1..16 */
  double mass;
  int gauge;         /* resident gauge field, 0 .. n_gauge-1 */
  uint64_t seed;     /* source seed */
  double arrival;    /* seconds after the start */
} gx_request;

typedef struct {
  int status;              /* GX_OK, GX_ENOCONV, GX_ENOMEM (larger than the whole budget), ... */
  int slot;
  int iterations;          /* largest over the right-hand sides, inner + patch-up */
  int outer_iterations;    /* defect-correction restarts */
  double residual;         /* largest |M x_k - b_k| / |b_k| recomputed in FP64 */
  double solver_residual;  /* largest true residual of the red-black CG (normal equations) */
  double t_admit, t_done;  /* seconds after the start */
  double footprint;        /* bytes charged to the budget when first placed on its slot */
  uint64_t digest;         /* gx_field_digest_slices of the request's FP64 solution slices */
  int batch;               /* requests in the solve that served it (1: a solve of its own) */
} gx_response;

typedef struct {
  int device;
  int slots;               /* concurrent solves: one stream and one host thread each */
  double budget;           /* bytes of device memory for request work (resident links excluded) */
  int n_gauge;             /* resident gauge fields per lattice */
  uint64_t gauge_seed;
  double gauge_eps;
  double tol;              /* FP64 true-residual target of every right-hand side */
  int mixed;               /* 1 (2): FP32 (FP16-storage) inner solves inside FP64 defect correction */
  int maxit;
  int realtime;
  int progress;            /* 0: silent; N > 0: one line on stderr after every N completed requests */
  const char *gauge_dir;   /* NULL: synthetic fields; else field g of lattice LxxLyxLzxLt is the NERSC
                            * file <gauge_dir>/ckpoint_lat.<Lx>x<Ly>x<Lz>x<Lt>.<g> (gx_nersc_read) */
  int fuse_rhs;            /* 0: every request is a solve of its own.  N > 0: when a request is admitted,
                            * waiting requests of the same lattice and gauge field join its solve (in
                            * arrival order) while the right-hand sides stay <= N and the footprint fits:
                            * one multi-RHS solve, each request's mass on its slices
                            * (gx_op_set_slice_masses).  Every request still reaches tol. */
} gx_serve_params;

/* Device bytes a request of nsrc sources on lattice dims holds while it runs (fields and the
 * operator's temporaries; links excluded). */
int gx_serve_footprint(const int dims[4], int nsrc, int mixed, double *bytes);
/* resp: n entries.  solutions: NULL, or n pointers (each NULL or room for the full host field of
 * the request's solution, layout of gx_field_export with Ls = nsrc).  wall: seconds from the start
 * to the last completion; resident: device bytes of the resident links. */
int gx_serve(const gx_serve_params *p, const gx_request *req, int n, gx_response *resp, double **solutions,
             double *wall, double *resident);

/* ---- launch-shape tuning (no Grid counterpart: Grid fixes its launch shape at build time,
 *      Grid/threads/Accelerator.h:416-461).  Process-wide knobs read by the kernel launchers;
 *      the autotuner in grid_amd/autotune.py sets them from measurements. ---- */
enum gx_tune_knob {
  GX_TUNE_DHOP_MAP = 0,     /* 0: linear workgroup->tile map, 1: XCD-contiguous slabs,
                               2, 3: XCDs take runs of 8 / 32 tiles of a common sweep,
                               4: blocked sweep (y-z blocks swept along t, sized for the L2), XCD slabs */
  GX_TUNE_DHOP_PF = 1,      /* 1: issue the next leg's neighbour loads before this leg's math */
  GX_TUNE_DHOP_NT = 2,      /* 1: non-temporal (streaming) output stores */
  GX_TUNE_DHOP_ORDER = 3,   /* 0: Grid's leg order, 1: the two legs of each direction adjacent */
  GX_TUNE_CG_GRAPH = 4,     /* 1: solver iteration batches are captured once and replayed as hipGraphs */
  GX_TUNE_FUSED = 5,        /* 1: Schur operator with fused hop epilogues (2-3 kernels per Mpc) */
  GX_TUNE_DHOP_SHAPE = 6,   /* hop tile shape: 0 = whole sites per wave, 4 waves (Ls <= 16); W > 0 = W waves of
                               64 consecutive (site, s) entries each, whole sites per workgroup (any Ls whose odd
                               part is at most 7, W a multiple of Ls / gcd(Ls, 64), W <= 8) */
  GX_TUNE_DHOP_PERSIST = 7, /* 1: the W-wave hop (GX_TUNE_DHOP_SHAPE > 0 or Ls > 16) runs as a persistent grid of
                               resident workgroups looping over the tiles */
  GX_TUNE_DHOP_MFMA = 8,    /* 1: FP32 multi-RHS hops (Ls = right-hand sides, 2..16, no matrix epilogue) do the link
                               products on the matrix cores (v_mfma_f32_16x16x4f32) */
  GX_TUNE_DHOP_TILE = 9,    /* 1: hops in tiles of 2 x 2 rows of x staged in LDS (x, y and z legs from the tile;
                               when 4 Lx/2 Ls threads and the tile fit a workgroup) */
  GX_TUNE_COUNT = 12
};
int gx_tune_set(int knob, int value);   /* every precision's value */
int gx_tune_get(int knob, int *value);  /* the FP64 value */
/* The knobs hold one value per storage precision (GX_F64, GX_F32, GX_F16): kernels on fields of a
 * precision take that precision's value, so the FP64 outer and the FP32 / FP16 inner operators of a
 * mixed-precision solve each run their own tuned launch shape. */
int gx_tune_set_prec(int knob, int prec, int value);
int gx_tune_get_prec(int knob, int prec, int *value);

/* ---- synthetic inputs (host utilities, no device work) ----
 * Gauge: per link G with Re,Im ~ N(0,1); H = (G+G^dag)/2 - tr/3; U = exp(i eps H) by
 * scaling and squaring of the Taylor series, projected back onto SU(3).  Gaussian numbers
 * from SplitMix64 counters keyed by (seed, global site, mu/s, element) and Box-Muller. */
int gx_random_gauge(const int dims[4], uint64_t seed, double eps, double *U);
int gx_gaussian_fermion(const int dims[4], int Ls, uint64_t seed, double *psi);   /* full field */
int gx_point_source(const int dims[4], int Ls, const int x[4], int spin, int colour, double *psi);
/* Device-side gx_gaussian_fermion straight into a full F64 or F32 field (same numbering; equal to
 * the host field up to the rounding of log, sqrt, sin and cos). */
int gx_field_gaussian(gx_field f, uint64_t seed);
/* No Grid counterpart: slices s0 .. s0 + ns - 1 of a full F64 or F32 field get the ns-slice field
 * gx_field_gaussian would give (several requests' sources side by side in one multi-RHS field). */
int gx_field_gaussian_slices(gx_field f, uint64_t seed, int s0, int ns);
/* Device-side gx_point_source into a full field: 1 at (x, spin, colour) in every slice s; with
 * spin = colour = -1, slice s (s < 12) gets spin-colour component s instead, the 12 point sources
 * of one site as one multi-RHS field. */
int gx_field_point_source(gx_field f, const int x[4], int spin, int colour);
/* Order-independent 64-bit digest of a field's stored bits: equal digests <=> (with overwhelming
 * probability) bitwise equal fields.  No Grid counterpart; used to compare solutions cheaply. */
int gx_field_digest(gx_field f, uint64_t *digest);
/* The same over the values of slices s0 .. s0 + ns - 1 of a full F64 field, indexed by their
 * logical position (site, slice - s0, component): a request's slices inside a wider multi-RHS
 * solution digest like the ns-slice field holding the same values. */
int gx_field_digest_slices(gx_field f, int s0, int ns, uint64_t *digest);

/* ---- gauge configurations ----
 * Grid: WilsonLoops<Gimpl>::avgPlaquette and linkTrace (WilsonLoops.h:121-126,203-219) of a host
 * gauge field (layout above), computed on the device: the average over sites and the 6 planes of
 * Re tr(U_mu(x) U_nu(x+mu) U_mu(x+nu)^dag U_nu(x)^dag) / 3, and over links of Re tr U / 3. */
int gx_gauge_statistics(gx_ctx ctx, const int dims[4], const double *U, double *plaquette, double *link_trace);

/* Header fields of a NERSC file as read, next to the values recomputed from its data. */
typedef struct {
  int dims[4];
  int two_row;                  /* DATATYPE 4D_SU3_GAUGE (1) or 4D_SU3_GAUGE_3x3 (0) */
  uint32_t checksum_header, checksum_data;
  double plaquette_header, plaquette;
  double link_trace_header, link_trace;
  int sequence_number;
  char floating_point[16];
  char ensemble_id[64];
  char ensemble_label[64];
} gx_nersc_info;

/* Grid: NerscIO::readConfiguration(Umu, header, file), NerscIO.h:63-221.  Reads a NERSC gauge file
 * (4D_SU3_GAUGE two-row or 4D_SU3_GAUGE_3x3; IEEE32BIG, IEEE32, IEEE64BIG, IEEE64/IEEE64LITTLE)
 * on the lattice dims into U (layout above).  As in Grid, it fails (GX_EINVAL, with the values in
 * the message and in *info) when the 32-bit word checksum differs from CHECKSUM, or when the
 * recomputed plaquette or link trace differs from the header by 1e-5 or 1e-6. */
int gx_nersc_read(gx_ctx ctx, const char *path, const int dims[4], double *U, gx_nersc_info *info);

/* Grid: NerscIO::writeConfiguration(Umu, file, two_row, bits32, ens_label, ens_id, sequence_number),
 * NerscIO.h:224-290.  Writes U as IEEE64BIG, three rows (two_row = 0) or two (two_row = 1), with
 * the plaquette, link trace and checksum in the header (bits32 has no effect in Grid either). */
int gx_nersc_write(gx_ctx ctx, const char *path, const int dims[4], const double *U, int two_row,
                   const char *ensemble_label, const char *ensemble_id, int sequence_number);

/* Header and metadata of an ILDG file as read, next to the values recomputed from its data. */
typedef struct {
  int dims[4];
  int precision;                          /* 32 or 64 (big-endian IEEE words) */
  uint32_t scidac_checksuma, scidac_checksumb;            /* computed from the binary record */
  uint32_t scidac_checksuma_file, scidac_checksumb_file;  /* scidac-checksum record */
  int has_plaquette;                      /* Grid grid-format or USQCD usqcdInfo metadata present */
  double plaquette_file, link_trace_file;
  double plaquette, link_trace;           /* recomputed on the device */
  int has_norm;                           /* Grid FieldNormMetaData record present */
  double norm2_file, norm2;
  char lfn[256];                          /* ildg-data-lfn */
} gx_ildg_info;

/* Grid: IldgReader::readConfiguration(Umu, FieldMetaData) (IldgIO.h:713-936): reads an ILDG (LIME)
 * gauge file on the lattice dims into U (layout above).  As in Grid it needs the ildg-format,
 * ildg-data-lfn, ildg-binary-data and scidac-checksum records, and fails (GX_EINVAL, values in
 * the message and *info) when the lattice or field type differs, the SciDAC checksums (BinaryIO.h:
 * 139-193) do not match, norm2 differs from a FieldNormMetaData record by 1e-5 (relative), or the
 * recomputed plaquette or link trace differs from Grid's or USQCD's metadata by 1e-5. */
int gx_ildg_read(gx_ctx ctx, const char *path, const int dims[4], double *U, gx_ildg_info *info);

/* Grid: IldgWriter::writeConfiguration(Umu, sequence, LFN, description) (IldgIO.h:633-690): writes U
 * as an ILDG file with big-endian IEEE words of `precision` bits (32 or 64): SciDAC file and record
 * metadata with the USQCD usqcdInfo (plaquette, link trace, description), ildg-format,
 * ildg-data-lfn (lfn), the binary record and its SciDAC checksums.  Grid's private grid-format and
 * FieldNormMetaData records are not written; readers then use the ILDG and USQCD records. */
int gx_ildg_write(gx_ctx ctx, const char *path, const int dims[4], const double *U, int precision, const char *lfn,
                  const char *description);

#ifdef __cplusplus
}
#endif
#endif /* GX_H */
