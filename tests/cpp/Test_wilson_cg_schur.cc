// Mirrors Grid's tests/solver/Test_wilson_cg_schur.cc: SchurRedBlackDiagMooeeSolve with
// ConjugateGradient on the Wilson operator; checks the unpreconditioned residual, and that the
// device-resident CG agrees with the same CG loop driven over the facade's BLAS1 (MdagM form).
#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian *UGrid = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()),
                                                        GridDefaultMpi());
  GridRedBlackCartesian *UrbGrid = SpaceTimeGrid::makeFourDimRedBlackGrid(UGrid);
  GridParallelRNG pRNG(UGrid);
  pRNG.SeedFixedIntegers(std::vector<int>({1, 2, 3, 4}));
  LatticeGaugeField Umu(UGrid);
  SU<Nc>::HotConfiguration(pRNG, Umu, 0.5);
  LatticeFermion src(UGrid); random(pRNG, src);
  LatticeFermion result(UGrid); result = Zero();
  LatticeFermion resid(UGrid);

  RealD mass = 0.1;
  WilsonFermionD Dw(Umu, *UGrid, *UrbGrid, mass);
  ConjugateGradient<LatticeFermion> CG(1.0e-10, 10000);
  SchurRedBlackDiagMooeeSolve<LatticeFermion> SchurSolver(CG);
  SchurSolver(Dw, src, result);
  Dw.M(result, resid);
  resid = resid - src;
  const double r = std::sqrt(norm2(resid) / norm2(src));
  std::cout << GridLogMessage << "unprec residual " << r << " iterations " << CG.IterationsToComplete << std::endl;
  int fail = !(r < 1e-8) || CG.TrueResidual > 1e-9;

  // the same Schur system through the generic host-driven loop (a user-defined LinearOperatorBase)
  struct Wrapped : public LinearOperatorBase<LatticeFermion> {
    SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &S;
    explicit Wrapped(SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &s) : S(s) {}
    void Op(const LatticeFermion &i, LatticeFermion &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermion &i, LatticeFermion &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermion &i, LatticeFermion &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermion &i, LatticeFermion &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  };
  SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> HermOpEO(Dw);
  Wrapped W(HermOpEO);
  LatticeFermion src_o(UrbGrid), x1(UrbGrid), x2(UrbGrid);
  pickCheckerboard(Odd, src_o, src);
  x1 = Zero();
  x2 = Zero();
  x1.Checkerboard() = Odd;
  x2.Checkerboard() = Odd;
  ConjugateGradient<LatticeFermion> CG1(1.0e-10, 10000), CG2(1.0e-10, 10000);
  CG1(HermOpEO, src_o, x1);
  CG2(W, src_o, x2);
  LatticeFermion d = x1 - x2;
  const double rel = std::sqrt(norm2(d) / norm2(x1));
  std::cout << GridLogMessage << "device CG " << CG1.IterationsToComplete << " its, host-driven CG "
            << CG2.IterationsToComplete << " its, relative difference " << rel << std::endl;
  fail |= std::abs(CG1.IterationsToComplete - CG2.IterationsToComplete) > 2 || !(rel < 1e-8);
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
