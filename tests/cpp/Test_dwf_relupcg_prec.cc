// Mirrors Grid's tests/solver/Test_dwf_relupcg_prec.cc: domain-wall (Shamir) Schur system solved
// with ConjugateGradientReliableUpdate (FP32 sloppy vectors, FP64 reliable updates, delta 0.2),
// with FP16-storage sloppy vectors, with FP16 sloppy vectors and a single-precision fallback
// (setFallbackLinop), and with plain FP64 ConjugateGradient; all reach the FP64 tolerance and agree.
#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  double relup_delta = 0.2;
  for (int i = 1; i < argc - 1; i++)
    if (std::string(argv[i]) == "--relup_delta") relup_delta = std::atof(argv[i + 1]);
  const int Ls = 12;
  GridCartesian *UGrid = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()),
                                                        GridDefaultMpi());
  GridRedBlackCartesian *UrbGrid = SpaceTimeGrid::makeFourDimRedBlackGrid(UGrid);
  GridCartesian *FGrid = SpaceTimeGrid::makeFiveDimGrid(Ls, UGrid);
  GridRedBlackCartesian *FrbGrid = SpaceTimeGrid::makeFiveDimRedBlackGrid(Ls, UGrid);
  GridParallelRNG RNG5(FGrid);
  RNG5.SeedFixedIntegers(std::vector<int>({5, 6, 7, 8}));
  GridParallelRNG RNG4(UGrid);
  RNG4.SeedFixedIntegers(std::vector<int>({1, 2, 3, 4}));
  LatticeFermionD src(FGrid); random(RNG5, src);
  LatticeGaugeFieldD Umu(UGrid);
  SU<Nc>::HotConfiguration(RNG4, Umu, 0.3);
  LatticeGaugeFieldF Umu_f(UGrid);
  precisionChange(Umu_f, Umu);

  RealD mass = 0.1, M5 = 1.8;
  DomainWallFermionD Ddwf(Umu, *FGrid, *FrbGrid, *UGrid, *UrbGrid, mass, M5);
  DomainWallFermionF Ddwf_f(Umu_f, *FGrid, *FrbGrid, *UGrid, *UrbGrid, mass, M5);
  LatticeFermionD src_o(FrbGrid), result_o(FrbGrid), result_h(FrbGrid), result_f(FrbGrid), result_o_2(FrbGrid);
  pickCheckerboard(Odd, src_o, src);
  result_o = Zero();
  result_o.Checkerboard() = Odd;
  result_h = Zero();
  result_h.Checkerboard() = Odd;
  result_f = Zero();
  result_f.Checkerboard() = Odd;
  result_o_2 = Zero();
  result_o_2.Checkerboard() = Odd;

  SchurDiagMooeeOperator<DomainWallFermionD, LatticeFermionD> HermOpEO(Ddwf);
  SchurDiagMooeeOperator<DomainWallFermionF, LatticeFermionF> HermOpEO_f(Ddwf_f);
  ConjugateGradientReliableUpdate<LatticeFermionD, LatticeFermionF> mCG(1e-10, 10000, relup_delta, FrbGrid,
                                                                        HermOpEO_f, HermOpEO);
  mCG(src_o, result_o);
  const int relups = mCG.ReliableUpdatesPerformed;
  const double res32 = mCG.TrueResidual;
  mCG.SloppyHalf = true;
  mCG(src_o, result_h);
  const double res16 = mCG.TrueResidual;
  mCG.setFallbackLinop(HermOpEO_f, 1e-6 * norm2(src_o));
  mCG(src_o, result_f);
  const double resfb = mCG.TrueResidual;
  ConjugateGradient<LatticeFermionD> CG(1.0e-10, 10000);
  CG(HermOpEO, src_o, result_o_2);
  LatticeFermionD diff = result_o - result_o_2, diffh = result_h - result_o_2, difff = result_f - result_o_2;
  const double nref = norm2(result_o_2);
  const double rel = std::sqrt(norm2(diff) / nref), relh = std::sqrt(norm2(diffh) / nref),
               relf = std::sqrt(norm2(difff) / nref);
  std::cout << GridLogMessage << "reliable-update true residual " << res32 << " (FP16 sloppy " << res16
            << ", FP16 with FP32 fallback " << resfb << "), FP64 true residual " << CG.TrueResidual
            << ", relative differences " << rel << " " << relh << " " << relf << ", reliable updates " << relups
            << std::endl;
  // user-defined operators (no device handle): Grid's reliable-update loop on the host over the facade
  struct PlainD : public LinearOperatorBase<LatticeFermionD> {
    SchurDiagMooeeOperator<DomainWallFermionD, LatticeFermionD> &S;
    explicit PlainD(SchurDiagMooeeOperator<DomainWallFermionD, LatticeFermionD> &s) : S(s) {}
    void Op(const LatticeFermionD &i, LatticeFermionD &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermionD &i, LatticeFermionD &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermionD &i, LatticeFermionD &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermionD &i, LatticeFermionD &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  } plain_d(HermOpEO);
  struct PlainF : public LinearOperatorBase<LatticeFermionF> {
    SchurDiagMooeeOperator<DomainWallFermionF, LatticeFermionF> &S;
    explicit PlainF(SchurDiagMooeeOperator<DomainWallFermionF, LatticeFermionF> &s) : S(s) {}
    void Op(const LatticeFermionF &i, LatticeFermionF &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermionF &i, LatticeFermionF &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermionF &i, LatticeFermionF &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermionF &i, LatticeFermionF &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  } plain_f(HermOpEO_f);
  ConjugateGradientReliableUpdate<LatticeFermionD, LatticeFermionF> hCG(1e-10, 10000, relup_delta, FrbGrid, plain_f,
                                                                        plain_d);
  LatticeFermionD result_p(FrbGrid);
  result_p = Zero();
  result_p.Checkerboard() = Odd;
  hCG(src_o, result_p);
  LatticeFermionD diffp = result_p - result_o_2;
  const double relp = std::sqrt(norm2(diffp) / nref);
  std::cout << GridLogMessage << "host-loop reliable update: " << hCG.IterationsToComplete << " its, "
            << hCG.ReliableUpdatesPerformed << " updates (device: " << relups << "), true residual "
            << hCG.TrueResidual << ", relative difference " << relp << std::endl;
  const int fail = !(res32 < 1e-9) || !(res16 < 1e-9) || !(resfb < 1e-9) || !(CG.TrueResidual < 1e-9) ||
                   !(rel < 1e-7) || !(relh < 1e-7) || !(relf < 1e-7) || relups < 1 || !(hCG.TrueResidual < 1e-9) ||
                   !(relp < 1e-7) || hCG.ReliableUpdatesPerformed < 1;
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
