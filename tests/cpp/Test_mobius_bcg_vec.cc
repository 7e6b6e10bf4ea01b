// Block CG over a list of Mobius fields through the Grid-API facade, in the three ways a Grid user
// can call BlockConjugateGradient(BlockCGrQVec) (Grid's tests/solver/Test_mobius_bcg_nosplit.cc calls
// it on MdagMLinearOperator):
//   unprec   MdagMLinearOperator on full 5D fields (device BlockCGrQVec, A = Mdag M);
//   schur    SchurDiagMooeeOperator on odd half fields (device BlockCGrQVec, A = MpcDagMpc);
//   custom   a user-defined LinearOperatorBase wrapping the Schur operator (the facade's own loop).
// Every block solution must match the ConjugateGradient solution of its source, and the block
// solver must need no more iterations than the slowest single solve.
#include <Grid/Grid.h>

#include <cstdio>

using namespace Grid;

namespace {

// A linear operator the facade knows nothing about: forwards to the Schur operator's MpcDagMpc.
template <class Field>
class Forwarding : public LinearOperatorBase<Field> {
 public:
  SchurOperatorBase<Field> &inner;
  explicit Forwarding(SchurOperatorBase<Field> &in) : inner(in) {}
  void Op(const Field &in, Field &out) override { inner.Mpc(in, out); }
  void AdjOp(const Field &in, Field &out) override { inner.MpcDag(in, out); }
  void HermOp(const Field &in, Field &out) override { inner.MpcDagMpc(in, out); }
  void HermOpAndNorm(const Field &in, Field &out, RealD &n1, RealD &n2) override {
    HermOp(in, out);
    n1 = innerProduct(in, out).real();
    n2 = norm2(out);
  }
};

int failures = 0;
void expect(bool ok, const char *what, double v) {
  std::printf("%-58s %.3e %s\n", what, v, ok ? "ok" : "FAIL");
  failures += ok ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  const int Ls = 6, nrhs = 3;
  const double tol = 1e-9;
  GridCartesian *UGrid = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()),
                                                        GridDefaultMpi());
  GridRedBlackCartesian *UrbGrid = SpaceTimeGrid::makeFourDimRedBlackGrid(UGrid);
  GridCartesian *FGrid = SpaceTimeGrid::makeFiveDimGrid(Ls, UGrid);
  GridRedBlackCartesian *FrbGrid = SpaceTimeGrid::makeFiveDimRedBlackGrid(Ls, UGrid);
  GridParallelRNG rng4(UGrid), rng5(FGrid);
  rng4.SeedFixedIntegers(std::vector<int>({31, 7, 2024, 5}));
  rng5.SeedFixedIntegers(std::vector<int>({11, 13, 17, 19}));
  LatticeGaugeFieldD U(UGrid);
  SU<Nc>::HotConfiguration(rng4, U, 0.3);
  MobiusFermionD D(U, *FGrid, *FrbGrid, *UGrid, *UrbGrid, 0.08, 1.8, 1.5, 0.5);

  std::vector<LatticeFermionD> src, src_o;
  for (int k = 0; k < nrhs; k++) {
    src.emplace_back(FGrid);
    random(rng5, src[k]);
    src_o.emplace_back(FrbGrid);
    pickCheckerboard(Odd, src_o[k], src[k]);
  }
  auto zeros = [&](GridBase *g, int cb) {
    std::vector<LatticeFermionD> v;
    for (int k = 0; k < nrhs; k++) {
      v.emplace_back(g);
      v[k] = Zero();
      if (g->checkerboarded) v[k].Checkerboard() = cb;
    }
    return v;
  };
  auto worst_diff = [&](const std::vector<LatticeFermionD> &a, const std::vector<LatticeFermionD> &b) {
    double w = 0;
    for (int k = 0; k < nrhs; k++) w = std::max(w, std::sqrt(norm2(a[k] - b[k]) / norm2(b[k])));
    return w;
  };

  // unpreconditioned normal equations
  {
    MdagMLinearOperator<MobiusFermionD, LatticeFermionD> mdagm(D);
    std::vector<LatticeFermionD> ref = zeros(FGrid, 0), blk = zeros(FGrid, 0);
    int slowest = 0;
    for (int k = 0; k < nrhs; k++) {
      ConjugateGradient<LatticeFermionD> cg(tol, 20000);
      cg(mdagm, src[k], ref[k]);
      slowest = std::max<int>(slowest, cg.IterationsToComplete);
    }
    BlockConjugateGradient<LatticeFermionD> bcg(BlockCGrQVec, 0, tol, 20000);
    bcg(mdagm, src, blk);
    expect(bcg.TrueResidual < 2 * tol, "unprec: block true residual", bcg.TrueResidual);
    expect(worst_diff(blk, ref) < 1e-6, "unprec: |x_block - x_cg| / |x_cg| (worst source)", worst_diff(blk, ref));
    expect(bcg.IterationsToComplete <= slowest + 1, "unprec: block iterations minus slowest CG",
           bcg.IterationsToComplete - slowest);
  }
  // Schur system, device block solver and the facade's loop on a foreign operator
  {
    SchurDiagMooeeOperator<MobiusFermionD, LatticeFermionD> schur(D);
    Forwarding<LatticeFermionD> custom(schur);
    std::vector<LatticeFermionD> ref = zeros(FrbGrid, Odd), dev = zeros(FrbGrid, Odd), host = zeros(FrbGrid, Odd);
    int slowest = 0;
    for (int k = 0; k < nrhs; k++) {
      ConjugateGradient<LatticeFermionD> cg(tol, 20000);
      cg(schur, src_o[k], ref[k]);
      slowest = std::max<int>(slowest, cg.IterationsToComplete);
    }
    BlockConjugateGradient<LatticeFermionD> bdev(BlockCGrQVec, 0, tol, 20000), bhost(BlockCGrQVec, 0, tol, 20000);
    bdev(schur, src_o, dev);
    bhost(custom, src_o, host);
    expect(bdev.TrueResidual < 2 * tol, "schur: device block true residual", bdev.TrueResidual);
    expect(bhost.TrueResidual < 2 * tol, "schur: facade-loop block true residual", bhost.TrueResidual);
    expect(worst_diff(dev, ref) < 1e-7, "schur: device block vs CG (worst source)", worst_diff(dev, ref));
    expect(worst_diff(host, dev) < 1e-9, "schur: facade loop vs device block", worst_diff(host, dev));
    expect(bdev.IterationsToComplete == bhost.IterationsToComplete, "schur: iterations device minus loop",
           bdev.IterationsToComplete - bhost.IterationsToComplete);
    expect(bdev.IterationsToComplete <= slowest + 1, "schur: block iterations minus slowest CG",
           bdev.IterationsToComplete - slowest);
  }
  Grid_finalize();
  std::printf("%s\n", failures ? "FAILED" : "PASSED");
  return failures ? 1 : 0;
}
