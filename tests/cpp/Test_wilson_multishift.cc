// Mirrors Grid's tests/solver/Test_staggered_multishift.cc with the Wilson Schur operator (staggered
// actions are not part of this library, and AlgRemez needs GMP, so the poles and residues are set by
// hand): ConjugateGradientMultiShift solves every shifted system at once; each solution agrees with
// ConjugateGradient on the shifted operator, driven through a user-defined LinearOperatorBase, and the
// rational sum psi = norm src + sum_s residues[s] x_s agrees with the sum of those solutions; on a
// user-defined operator the facade runs the same loop on the host with the same results.
#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian Grid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridRedBlackCartesian RBGrid(&Grid);
  GridParallelRNG pRNG(&Grid);
  pRNG.SeedFixedIntegers(std::vector<int>({1, 2, 3, 4}));
  LatticeGaugeField Umu(&Grid);
  SU<Nc>::HotConfiguration(pRNG, Umu, 0.5);
  LatticeFermion src(&Grid);
  random(pRNG, src);
  LatticeFermion src_o(&RBGrid);
  pickCheckerboard(Odd, src_o, src);

  const int degree = 4;
  MultiShiftFunction Shifts(degree, 0.01, 10.0);
  const double poles[degree] = {0.01, 0.05, 0.3, 2.0}, residues[degree] = {0.4, -0.2, 0.7, 1.1};
  for (int s = 0; s < degree; s++) {
    Shifts.poles[s] = poles[s];
    Shifts.residues[s] = residues[s];
    Shifts.tolerances[s] = 1.0e-10;
  }
  Shifts.norm = 0.25;

  RealD mass = 0.1;
  WilsonFermionD Dw(Umu, Grid, RBGrid, mass);
  SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> HermOpEO(Dw);
  std::vector<LatticeFermion> result;
  for (int s = 0; s < degree; s++) result.emplace_back(&RBGrid);
  LatticeFermion psi(&RBGrid);
  ConjugateGradientMultiShift<LatticeFermion> MSCG(10000, Shifts);
  MSCG(HermOpEO, src_o, result, psi);

  struct Shifted : public LinearOperatorBase<LatticeFermion> {
    SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &S;
    RealD sigma;
    Shifted(SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &s, RealD sg) : S(s), sigma(sg) {}
    void Op(const LatticeFermion &i, LatticeFermion &o) override { HermOp(i, o); }
    void AdjOp(const LatticeFermion &i, LatticeFermion &o) override { HermOp(i, o); }
    void HermOp(const LatticeFermion &i, LatticeFermion &o) override {
      S.MpcDagMpc(i, o);
      o = o + sigma * i;
    }
    void HermOpAndNorm(const LatticeFermion &i, LatticeFermion &o, RealD &a, RealD &b) override {
      HermOp(i, o);
      a = innerProduct(i, o).real();
      b = norm2(o);
    }
  };
  int fail = 0;
  LatticeFermion sum(&RBGrid);
  sum = Shifts.norm * src_o;
  for (int s = 0; s < degree; s++) {
    Shifted A(HermOpEO, poles[s]);
    LatticeFermion x(&RBGrid);
    x = Zero();
    x.Checkerboard() = Odd;
    ConjugateGradient<LatticeFermion> CG(1.0e-10, 10000);
    CG(A, src_o, x);
    LatticeFermion d = x - result[s];
    const double rel = std::sqrt(norm2(d) / norm2(x));
    std::cout << GridLogMessage << "shift " << poles[s] << ": multishift " << MSCG.IterationsToCompleteShift[s]
              << " its, true residual " << MSCG.TrueResidualShift[s] << "; single-shift CG " << CG.IterationsToComplete
              << " its; relative difference " << rel << std::endl;
    fail |= !(MSCG.TrueResidualShift[s] < 1e-9) || !(rel < 1e-8);
    sum = sum + residues[s] * x;
  }
  LatticeFermion dd = sum - psi;
  const double rsum = std::sqrt(norm2(dd) / norm2(sum));
  std::cout << GridLogMessage << "rational sum relative difference " << rsum << std::endl;
  fail |= !(rsum < 1e-8);

  // a user-defined operator (no device handle): Grid's loop over the facade BLAS gives the same solutions
  struct Plain : public LinearOperatorBase<LatticeFermion> {
    SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &S;
    explicit Plain(SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> &s) : S(s) {}
    void Op(const LatticeFermion &i, LatticeFermion &o) override { S.Mpc(i, o); }
    void AdjOp(const LatticeFermion &i, LatticeFermion &o) override { S.MpcDag(i, o); }
    void HermOp(const LatticeFermion &i, LatticeFermion &o) override { S.MpcDagMpc(i, o); }
    void HermOpAndNorm(const LatticeFermion &i, LatticeFermion &o, RealD &a, RealD &b) override {
      S.HermOpAndNorm(i, o, a, b);
    }
  } plain(HermOpEO);
  std::vector<LatticeFermion> result2;
  for (int s = 0; s < degree; s++) result2.emplace_back(&RBGrid);
  ConjugateGradientMultiShift<LatticeFermion> MSCG2(10000, Shifts);
  MSCG2(plain, src_o, result2);
  double worst = 0;
  for (int s = 0; s < degree; s++) {
    LatticeFermion e = result2[s] - result[s];
    worst = std::max(worst, std::sqrt(norm2(e) / norm2(result[s])));
    fail |= MSCG2.IterationsToCompleteShift[s] != MSCG.IterationsToCompleteShift[s] || !(MSCG2.TrueResidualShift[s] < 1e-9);
  }
  std::cout << GridLogMessage << "host-loop multishift vs device multishift: " << worst << std::endl;
  fail |= !(worst < 1e-10);
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
