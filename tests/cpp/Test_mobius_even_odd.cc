// Mirrors Grid's tests/core/Test_mobius_even_odd.cc on the gx facade: Meo + Moe + Moo + Mee = M,
// adjointness of Meooe/Mooee, MooeeInv Mooee = 1, MpcDagMpc Hermitian, for b=1.5, c=0.5, M5=1.8.
#include <Grid/Grid.h>

using namespace Grid;

static int failures = 0;
static void expect_small(const char *what, double v, double tol) {
  std::cout << GridLogMessage << what << " = " << v << (v < tol ? "  ok" : "  FAIL") << std::endl;
  if (!(v < tol)) failures++;
}

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

  LatticeFermion src(FGrid); random(RNG5, src);
  LatticeFermion phi(FGrid); random(RNG5, phi);
  LatticeFermion chi(FGrid); random(RNG5, chi);
  LatticeFermion result(FGrid), ref(FGrid);
  LatticeGaugeField Umu(UGrid);
  SU<Nc>::HotConfiguration(RNG4, Umu);

  RealD mass = 0.1, M5 = 1.8, b = 1.5, c = 0.5;
  MobiusFermionD Ddwf(Umu, *FGrid, *FrbGrid, *UGrid, *UrbGrid, mass, M5, b, c);

  LatticeFermion src_e(FrbGrid), src_o(FrbGrid), r_e(FrbGrid), r_o(FrbGrid), r_eo(FGrid);
  std::cout << GridLogMessage << "= Testing that Meo + Moe + Moo + Mee = Munprec" << std::endl;
  pickCheckerboard(Even, src_e, src);
  pickCheckerboard(Odd, src_o, src);
  Ddwf.Meooe(src_e, r_o);
  Ddwf.Meooe(src_o, r_e);
  setCheckerboard(r_eo, r_o);
  setCheckerboard(r_eo, r_e);
  Ddwf.Mooee(src_e, r_e);
  Ddwf.Mooee(src_o, r_o);
  LatticeFermion r_d(FGrid);
  setCheckerboard(r_d, r_e);
  setCheckerboard(r_d, r_o);
  r_eo = r_eo + r_d;
  Ddwf.M(src, ref);
  LatticeFermion err = ref - r_eo;
  expect_small("EO norm diff / norm", norm2(err) / norm2(ref), 1e-28);

  LatticeFermion chi_e(FrbGrid), chi_o(FrbGrid), dchi_e(FrbGrid), dchi_o(FrbGrid);
  LatticeFermion phi_e(FrbGrid), phi_o(FrbGrid), dphi_e(FrbGrid), dphi_o(FrbGrid);
  pickCheckerboard(Even, chi_e, chi);
  pickCheckerboard(Odd, chi_o, chi);
  pickCheckerboard(Even, phi_e, phi);
  pickCheckerboard(Odd, phi_o, phi);
  std::cout << GridLogMessage << "= Test Ddagger is the dagger of D: Meooe" << std::endl;
  Ddwf.Meooe(chi_e, dchi_o);
  Ddwf.Meooe(chi_o, dchi_e);
  Ddwf.MeooeDag(phi_e, dphi_o);
  Ddwf.MeooeDag(phi_o, dphi_e);
  ComplexD pDce = innerProduct(phi_e, dchi_e), pDco = innerProduct(phi_o, dchi_o);
  ComplexD cDpe = innerProduct(chi_e, dphi_e), cDpo = innerProduct(chi_o, dphi_o);
  expect_small("Meooe e adjoint", std::abs(pDce - std::conj(cDpo)) / std::abs(pDce), 1e-13);
  expect_small("Meooe o adjoint", std::abs(pDco - std::conj(cDpe)) / std::abs(pDco), 1e-13);
  std::cout << GridLogMessage << "= Test Ddagger is the dagger of D: Mooee" << std::endl;
  Ddwf.Mooee(chi_e, dchi_e);
  Ddwf.MooeeDag(phi_e, dphi_e);
  pDce = innerProduct(phi_e, dchi_e);
  cDpe = innerProduct(chi_e, dphi_e);
  expect_small("Mooee adjoint", std::abs(pDce - std::conj(cDpe)) / std::abs(pDce), 1e-13);
  std::cout << GridLogMessage << "= Test MeeInv Mee = 1" << std::endl;
  Ddwf.Mooee(chi_e, src_e);
  Ddwf.MooeeInv(src_e, phi_e);
  err = phi_e - chi_e;
  expect_small("MooeeInv Mooee - 1", norm2(err) / norm2(chi_e), 1e-28);
  Ddwf.MooeeDag(chi_o, src_o);
  Ddwf.MooeeInvDag(src_o, phi_o);
  err = phi_o - chi_o;
  expect_small("MooeeInvDag MooeeDag - 1", norm2(err) / norm2(chi_o), 1e-28);
  std::cout << GridLogMessage << "= Test MpcDagMpc is Hermitian" << std::endl;
  random(RNG5, phi);
  random(RNG5, chi);
  pickCheckerboard(Odd, chi_o, chi);
  pickCheckerboard(Odd, phi_o, phi);
  SchurDiagMooeeOperator<MobiusFermionD, LatticeFermion> HermOpEO(Ddwf);
  HermOpEO.MpcDagMpc(chi_o, dchi_o);
  HermOpEO.MpcDagMpc(phi_o, dphi_o);
  pDco = innerProduct(phi_o, dchi_o);
  cDpo = innerProduct(chi_o, dphi_o);
  expect_small("o hermiticity", std::abs(pDco - std::conj(cDpo)) / std::abs(pDco), 1e-13);

  Grid_finalize();
  std::cout << (failures ? "FAILED" : "PASSED") << std::endl;
  return failures ? 1 : 0;
}
