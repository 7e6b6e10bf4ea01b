// Mirrors Grid's tests/Test_dwf_mixedcg_prec.cc: Mobius DWF Schur system solved with
// MixedPrecisionConjugateGradient (FP32 inner, then FP16-storage inner, FP64 outer) and with plain
// FP64 ConjugateGradient; all reach the FP64 tolerance and agree.
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
  GridParallelRNG RNG5(FGrid);
  RNG5.SeedFixedIntegers(std::vector<int>({1, 2, 3, 4}));
  GridParallelRNG RNG4(UGrid);
  RNG4.SeedFixedIntegers(std::vector<int>({5, 6, 7, 8}));
  LatticeFermionD src(FGrid); random(RNG5, src);
  LatticeGaugeFieldD Umu(UGrid);
  SU<Nc>::HotConfiguration(RNG4, Umu, 0.3);
  LatticeGaugeFieldF Umu_f(UGrid);
  precisionChange(Umu_f, Umu);

  RealD mass = 0.1, M5 = 1.8, b = 1.5, c = 0.5;
  MobiusFermionD Ddwf(Umu, *FGrid, *FrbGrid, *UGrid, *UrbGrid, mass, M5, b, c);
  MobiusFermionF Ddwf_f(Umu_f, *FGrid, *FrbGrid, *UGrid, *UrbGrid, mass, M5, b, c);
  LatticeFermionD src_o(FrbGrid), result_o(FrbGrid), result_o_2(FrbGrid);
  pickCheckerboard(Odd, src_o, src);
  result_o = Zero();
  result_o.Checkerboard() = Odd;
  result_o_2 = Zero();
  result_o_2.Checkerboard() = Odd;

  SchurDiagMooeeOperator<MobiusFermionD, LatticeFermionD> HermOpEO(Ddwf);
  SchurDiagMooeeOperator<MobiusFermionF, LatticeFermionF> HermOpEO_f(Ddwf_f);
  MixedPrecisionConjugateGradient<LatticeFermionD, LatticeFermionF> mCG(1.0e-10, 10000, 50, FrbGrid, HermOpEO_f,
                                                                         HermOpEO);
  mCG.InnerTolerance = 1.0e-5;
  mCG(src_o, result_o);
  ConjugateGradient<LatticeFermionD> CG(1.0e-10, 10000);
  CG(HermOpEO, src_o, result_o_2);
  LatticeFermionD diff = result_o - result_o_2;
  const double rel = std::sqrt(norm2(diff) / norm2(result_o_2));
  // the same solve with FP16-storage inner solves
  LatticeFermionD result_h(FrbGrid);
  result_h = Zero();
  result_h.Checkerboard() = Odd;
  mCG.InnerHalf = true;
  mCG(src_o, result_h);
  LatticeFermionD diffh = result_h - result_o_2;
  const double relh = std::sqrt(norm2(diffh) / norm2(result_o_2));
  // user-defined operators (no device handle): Grid's mixed loop on the host over the facade
  struct PlainD : public LinearOperatorBase<LatticeFermionD> {
    SchurDiagMooeeOperator<MobiusFermionD, LatticeFermionD> &S;
    explicit PlainD(SchurDiagMooeeOperator<MobiusFermionD, LatticeFermionD> &s) : S(s) {}
    void Op(const LatticeFermionD &i, LatticeFermionD &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermionD &i, LatticeFermionD &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermionD &i, LatticeFermionD &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermionD &i, LatticeFermionD &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  } plain_d(HermOpEO);
  struct PlainF : public LinearOperatorBase<LatticeFermionF> {
    SchurDiagMooeeOperator<MobiusFermionF, LatticeFermionF> &S;
    explicit PlainF(SchurDiagMooeeOperator<MobiusFermionF, LatticeFermionF> &s) : S(s) {}
    void Op(const LatticeFermionF &i, LatticeFermionF &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermionF &i, LatticeFermionF &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermionF &i, LatticeFermionF &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermionF &i, LatticeFermionF &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  } plain_f(HermOpEO_f);
  MixedPrecisionConjugateGradient<LatticeFermionD, LatticeFermionF> hCG(1.0e-10, 10000, 50, FrbGrid, plain_f, plain_d);
  hCG.InnerTolerance = 1.0e-5;
  LatticeFermionD result_p(FrbGrid);
  result_p = Zero();
  result_p.Checkerboard() = Odd;
  hCG(src_o, result_p);
  LatticeFermionD diffp = result_p - result_o_2;
  const double relp = std::sqrt(norm2(diffp) / norm2(result_o_2));
  std::cout << GridLogMessage << "mixed true residual " << mCG.TrueResidual << " (FP16 inner), host-loop mixed "
            << hCG.TrueResidual << " (" << hCG.TotalOuterIterations << " restarts), FP64 true residual "
            << CG.TrueResidual << ", relative differences " << rel << " " << relh << " " << relp << std::endl;
  const int fail = !(mCG.TrueResidual < 1e-9) || !(CG.TrueResidual < 1e-9) || !(rel < 1e-7) || !(relh < 1e-7) ||
                   !(hCG.TrueResidual < 1e-9) || !(relp < 1e-7) || hCG.TotalOuterIterations < 1;
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
