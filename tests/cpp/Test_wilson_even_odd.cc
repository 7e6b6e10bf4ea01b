// Mirrors Grid's tests/core/Test_wilson_even_odd.cc on the gx facade: Deo + Doe = Dunprec,
// <phi|Deo chi>^* = <chi|Deo^dag phi>, MooeeInv Mooee = 1 (and Dag), MpcDagMpc Hermitian.
#include <Grid/Grid.h>

using namespace Grid;

static int failures = 0;
static void expect_small(const char *what, double v, double tol) {
  std::cout << GridLogMessage << what << " = " << v << (v < tol ? "  ok" : "  FAIL") << std::endl;
  if (!(v < tol)) failures++;
}

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  Coordinate latt_size = GridDefaultLatt();
  GridCartesian Grid(latt_size, GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridRedBlackCartesian RBGrid(&Grid);
  GridParallelRNG pRNG(&Grid);
  pRNG.SeedFixedIntegers(std::vector<int>({45, 12, 81, 9}));

  LatticeFermion src(&Grid); random(pRNG, src);
  LatticeFermion phi(&Grid); random(pRNG, phi);
  LatticeFermion chi(&Grid); random(pRNG, chi);
  LatticeFermion ref(&Grid); ref = Zero();
  LatticeGaugeField Umu(&Grid);
  SU<Nc>::HotConfiguration(pRNG, Umu);

  RealD mass = 0.1;
  WilsonFermionD Dw(Umu, Grid, RBGrid, mass);

  LatticeFermion src_e(&RBGrid), src_o(&RBGrid), r_e(&RBGrid), r_o(&RBGrid), r_eo(&Grid);
  std::cout << GridLogMessage << "= Testing that Deo + Doe = Dunprec" << std::endl;
  pickCheckerboard(Even, src_e, src);
  pickCheckerboard(Odd, src_o, src);
  Dw.Meooe(src_e, r_o);
  Dw.Meooe(src_o, r_e);
  Dw.Dhop(src, ref, DaggerNo);
  setCheckerboard(r_eo, r_o);
  setCheckerboard(r_eo, r_e);
  LatticeFermion err = ref - r_eo;
  expect_small("EO norm diff / norm", norm2(err) / norm2(ref), 1e-28);

  std::cout << GridLogMessage << "= Test Ddagger is the dagger of D: <phi|Deo|chi>* = <chi|Deo^dag|phi>" << std::endl;
  LatticeFermion chi_e(&RBGrid), chi_o(&RBGrid), dchi_e(&RBGrid), dchi_o(&RBGrid);
  LatticeFermion phi_e(&RBGrid), phi_o(&RBGrid), dphi_e(&RBGrid), dphi_o(&RBGrid);
  pickCheckerboard(Even, chi_e, chi);
  pickCheckerboard(Odd, chi_o, chi);
  pickCheckerboard(Even, phi_e, phi);
  pickCheckerboard(Odd, phi_o, phi);
  Dw.Meooe(chi_e, dchi_o);
  Dw.Meooe(chi_o, dchi_e);
  Dw.MeooeDag(phi_e, dphi_o);
  Dw.MeooeDag(phi_o, dphi_e);
  ComplexD pDce = innerProduct(phi_e, dchi_e);
  ComplexD pDco = innerProduct(phi_o, dchi_o);
  ComplexD cDpe = innerProduct(chi_e, dphi_e);
  ComplexD cDpo = innerProduct(chi_o, dphi_o);
  expect_small("e adjoint mismatch", std::abs(pDce - std::conj(cDpo)) / std::abs(pDce), 1e-13);
  expect_small("o adjoint mismatch", std::abs(pDco - std::conj(cDpe)) / std::abs(pDco), 1e-13);

  std::cout << GridLogMessage << "= Test MeeInv Mee = 1" << std::endl;
  pickCheckerboard(Even, chi_e, chi);
  pickCheckerboard(Odd, chi_o, chi);
  Dw.Mooee(chi_e, src_e);
  Dw.MooeeInv(src_e, phi_e);
  Dw.Mooee(chi_o, src_o);
  Dw.MooeeInv(src_o, phi_o);
  setCheckerboard(phi, phi_e);
  setCheckerboard(phi, phi_o);
  err = phi - chi;
  expect_small("MooeeInv Mooee - 1", norm2(err) / norm2(chi), 1e-28);
  Dw.MooeeDag(chi_e, src_e);
  Dw.MooeeInvDag(src_e, phi_e);
  Dw.MooeeDag(chi_o, src_o);
  Dw.MooeeInvDag(src_o, phi_o);
  setCheckerboard(phi, phi_e);
  setCheckerboard(phi, phi_o);
  err = phi - chi;
  expect_small("MooeeInvDag MooeeDag - 1", norm2(err) / norm2(chi), 1e-28);

  std::cout << GridLogMessage << "= Test MpcDagMpc is Hermitian" << std::endl;
  random(pRNG, phi);
  random(pRNG, chi);
  pickCheckerboard(Even, chi_e, chi);
  pickCheckerboard(Odd, chi_o, chi);
  pickCheckerboard(Even, phi_e, phi);
  pickCheckerboard(Odd, phi_o, phi);
  SchurDiagMooeeOperator<WilsonFermionD, LatticeFermion> HermOpEO(Dw);
  HermOpEO.MpcDagMpc(chi_e, dchi_e);
  HermOpEO.MpcDagMpc(chi_o, dchi_o);
  HermOpEO.MpcDagMpc(phi_e, dphi_e);
  HermOpEO.MpcDagMpc(phi_o, dphi_o);
  pDce = innerProduct(phi_e, dchi_e);
  pDco = innerProduct(phi_o, dchi_o);
  cDpe = innerProduct(chi_e, dphi_e);
  cDpo = innerProduct(chi_o, dphi_o);
  expect_small("e hermiticity", std::abs(pDce - std::conj(cDpe)) / std::abs(pDce), 1e-13);
  expect_small("o hermiticity", std::abs(pDco - std::conj(cDpo)) / std::abs(pDco), 1e-13);

  Grid_finalize();
  std::cout << (failures ? "FAILED" : "PASSED") << std::endl;
  return failures ? 1 : 0;
}
