// The single-direction hop and the force through the Grid API: the eight DhopDir legs sum to Dhop,
// MdirAll returns them in Grid's order, and for every mu the trace of MDeriv's colour matrices summed over the lattice is <A, Mdir(B, mu, +1)>
// (InsertForce5D's outer product traced over colour).
#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian Grid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridRedBlackCartesian RBGrid(&Grid);
  GridParallelRNG pRNG(&Grid);
  pRNG.SeedFixedIntegers(std::vector<int>({5, 6, 7, 8}));
  LatticeGaugeField Umu(&Grid);
  SU<Nc>::HotConfiguration(pRNG, Umu, 0.5);
  LatticeFermion A(&Grid), B(&Grid), out(&Grid), sum(&Grid), ref(&Grid);
  random(pRNG, A);
  random(pRNG, B);
  WilsonFermionD Dw(Umu, Grid, RBGrid, 0.1);
  sum = Zero();
  for (int mu = 0; mu < Nd; mu++)
    for (int disp = -1; disp <= 1; disp += 2) {
      Dw.DhopDir(B, out, mu, disp);
      sum = sum + out;
    }
  Dw.Dhop(B, ref, DaggerNo);
  LatticeFermion d = sum - ref;
  const double rsum = std::sqrt(norm2(d) / norm2(ref));
  // MdirAll gives the same eight legs, indexed dir + 4 (disp = -1)
  std::vector<LatticeFermion> all;
  for (int k = 0; k < 8; k++) all.emplace_back(&Grid);
  Dw.MdirAll(B, all);
  double rall = 0;
  for (int dir = 0; dir < Nd; dir++)
    for (int disp = -1; disp <= 1; disp += 2) {
      Dw.Mdir(B, out, dir, disp);
      LatticeFermion e = out - all[dir + (disp == 1 ? 0 : 4)];
      rall = std::max(rall, norm2(e));
    }
  LatticeGaugeField mat(&Grid);
  Dw.MDeriv(mat, A, B, DaggerNo);
  double worst = 0;
  const size_t V = mat.data.size() / 36;
  for (int mu = 0; mu < Nd; mu++) {
    Dw.Mdir(B, out, mu, 1);
    const ComplexD ip = innerProduct(A, out);
    ComplexD tr = 0;
    for (size_t x = 0; x < V; x++)
      for (int a = 0; a < 3; a++) tr += mat.data[(x * 4 + mu) * 9 + a * 3 + a];
    worst = std::max(worst, std::abs(tr - ip) / std::abs(ip));
  }
  std::cout << GridLogMessage << "sum of DhopDir legs vs Dhop: " << rsum << "; trace of MDeriv vs <A, Mdir B>: " << worst
            << std::endl;
  std::cout << GridLogMessage << "MdirAll vs Mdir: " << rall << std::endl;
  const int fail = !(rsum < 1e-14) || !(worst < 1e-12) || rall != 0.0;
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
