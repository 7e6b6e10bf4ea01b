// Mirrors Grid's tests/solver/Test_staggered_block_cg_prec.cc with the Wilson operator (staggered
// actions are not part of this library): Ls = 8 right-hand sides along the fifth dimension of a
// multi-RHS Wilson operator, solved on the Schur system as one 5D ConjugateGradient, with
// BlockConjugateGradient(BlockCGrQ) and with BlockConjugateGradient(CGmultiRHS); all reach the
// tolerance and agree, and the block Krylov solver needs no more iterations than CGmultiRHS.
#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  const int Ls = 8;
  GridCartesian *UGrid = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()),
                                                        GridDefaultMpi());
  GridRedBlackCartesian *UrbGrid = SpaceTimeGrid::makeFourDimRedBlackGrid(UGrid);
  GridCartesian *FGrid = SpaceTimeGrid::makeFiveDimGrid(Ls, UGrid);
  GridRedBlackCartesian *FrbGrid = SpaceTimeGrid::makeFiveDimRedBlackGrid(Ls, UGrid);
  (void)UrbGrid;
  std::vector<int> seeds({1, 2, 3, 4});
  GridParallelRNG pRNG(UGrid);
  pRNG.SeedFixedIntegers(seeds);
  GridParallelRNG pRNG5(FGrid);
  pRNG5.SeedFixedIntegers(seeds);
  LatticeFermion src(FGrid);
  random(pRNG5, src);
  LatticeFermion src_o(FrbGrid);
  pickCheckerboard(Odd, src_o, src);
  LatticeGaugeField Umu(UGrid);
  SU<Nc>::HotConfiguration(pRNG, Umu, 0.5);

  RealD mass = 0.1;
  WilsonFermionD Dw(Umu, *FGrid, *FrbGrid, mass);
  SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> HermOp(Dw);
  ConjugateGradient<LatticeFermion> CG(1.0e-10, 10000);
  BlockConjugateGradient<LatticeFermion> BCGrQ(BlockCGrQ, 0, 1.0e-10, 10000);
  BlockConjugateGradient<LatticeFermion> mCG(CGmultiRHS, 0, 1.0e-10, 10000);

  LatticeFermion x_cg(FrbGrid), x_rq(FrbGrid), x_m(FrbGrid);
  x_cg = Zero();
  x_rq = Zero();
  x_m = Zero();
  x_cg.Checkerboard() = x_rq.Checkerboard() = x_m.Checkerboard() = Odd;
  CG(HermOp, src_o, x_cg);
  BCGrQ(HermOp, src_o, x_rq);
  mCG(HermOp, src_o, x_m);
  LatticeFermion d1 = x_rq - x_cg, d2 = x_m - x_cg;
  const double n = norm2(x_cg);
  const double r1 = std::sqrt(norm2(d1) / n), r2 = std::sqrt(norm2(d2) / n);
  std::cout << GridLogMessage << "5D CG " << CG.IterationsToComplete << " its (true residual " << CG.TrueResidual
            << "); BlockCGrQ " << BCGrQ.IterationsToComplete << " its (" << BCGrQ.TrueResidual << ", diff " << r1
            << "); CGmultiRHS " << mCG.IterationsToComplete << " its (" << mCG.TrueResidual << ", diff " << r2 << ")"
            << std::endl;
  const int fail = !(CG.TrueResidual < 1e-9) || !(BCGrQ.TrueResidual < 1e-9) || !(mCG.TrueResidual < 1e-9) ||
                   !(r1 < 1e-8) || !(r2 < 1e-8) || BCGrQ.IterationsToComplete > mCG.IterationsToComplete ||
                   BCGrQ.Nblock != Ls;
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
