// MixedPrecisionConjugateGradientBatched through the Grid-API facade: a batch of odd-checkerboard
// Schur sources solved in FP32-inner defect correction must give each source the solution the
// FP64 ConjugateGradient gives it.  Three paths:
//   wilson   single-RHS Wilson action: the device batch runs as one multi-RHS solve;
//   mobius   Mobius action: the device batch runs source by source;
//   loop     updateResidual = false: Grid's loop in the facade.
#include <Grid/Grid.h>

#include <cstdio>

using namespace Grid;

namespace {

int failures = 0;
void expect(bool ok, const char *what, double v) {
  std::printf("%-58s %.3e %s\n", what, v, ok ? "ok" : "FAIL");
  failures += ok ? 0 : 1;
}

template <class ActionD, class ActionF>
void run(const char *name, ActionD &Dd, ActionF &Df, GridBase *full, GridBase *half, GridParallelRNG &rng, int nsrc,
         bool update_residual) {
  SchurDiagMooeeOperator<ActionD, LatticeFermionD> opd(Dd);
  SchurDiagMooeeOperator<ActionF, LatticeFermionF> opf(Df);
  const double tol = 1e-10;
  std::vector<LatticeFermionD> src, ref, sol;
  for (int k = 0; k < nsrc; k++) {
    LatticeFermionD s(full);
    random(rng, s);
    src.emplace_back(half);
    pickCheckerboard(Odd, src[k], s);
    ref.emplace_back(half);
    sol.emplace_back(half);
    ref[k] = Zero();
    sol[k] = Zero();
    ref[k].Checkerboard() = sol[k].Checkerboard() = Odd;
    ConjugateGradient<LatticeFermionD> cg(tol, 20000);
    cg(opd, src[k], ref[k]);
  }
  MixedPrecisionConjugateGradientBatched<LatticeFermionD, LatticeFermionF> mcg(tol, 20000, 50, 20000, half, opf, opd,
                                                                               update_residual);
  mcg.InnerTolerance = 1e-5;
  mcg(src, sol);
  double worst = 0, worst_res = 0;
  for (int k = 0; k < nsrc; k++) {
    worst = std::max(worst, std::sqrt(norm2(sol[k] - ref[k]) / norm2(ref[k])));
    LatticeFermionD a(half);
    a.Checkerboard() = Odd;
    opd.HermOp(sol[k], a);
    worst_res = std::max(worst_res, std::sqrt(norm2(a - src[k]) / norm2(src[k])));
  }
  std::string tag(name);
  expect(worst_res < 2 * tol, (tag + ": worst |A x - b| / |b|").c_str(), worst_res);
  expect(worst < 1e-8, (tag + ": worst |x_batched - x_cg| / |x_cg|").c_str(), worst);
}

}  // namespace

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian *UGrid = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()),
                                                        GridDefaultMpi());
  GridRedBlackCartesian *UrbGrid = SpaceTimeGrid::makeFourDimRedBlackGrid(UGrid);
  GridParallelRNG rng(UGrid);
  rng.SeedFixedIntegers(std::vector<int>({88, 1, 512, 3}));
  LatticeGaugeFieldD U(UGrid);
  SU<Nc>::HotConfiguration(rng, U, 0.4);
  LatticeGaugeFieldF Uf(UGrid);
  precisionChange(Uf, U);
  {
    WilsonFermionD Dd(U, *UGrid, *UrbGrid, 0.07);
    WilsonFermionF Df(Uf, *UGrid, *UrbGrid, 0.07);
    run("wilson", Dd, Df, UGrid, UrbGrid, rng, 5, true);
    run("loop", Dd, Df, UGrid, UrbGrid, rng, 2, false);
  }
  {
    const int Ls = 6;
    GridCartesian *FGrid = SpaceTimeGrid::makeFiveDimGrid(Ls, UGrid);
    GridRedBlackCartesian *FrbGrid = SpaceTimeGrid::makeFiveDimRedBlackGrid(Ls, UGrid);
    GridParallelRNG rng5(FGrid);
    rng5.SeedFixedIntegers(std::vector<int>({4, 3, 2, 1}));
    MobiusFermionD Dd(U, *FGrid, *FrbGrid, *UGrid, *UrbGrid, 0.05, 1.8, 1.5, 0.5);
    MobiusFermionF Df(Uf, *FGrid, *FrbGrid, *UGrid, *UrbGrid, 0.05, 1.8, 1.5, 0.5);
    run("mobius", Dd, Df, FGrid, FrbGrid, rng5, 2, true);
  }
  Grid_finalize();
  std::printf("%s\n", failures ? "FAILED" : "PASSED");
  return failures ? 1 : 0;
}
