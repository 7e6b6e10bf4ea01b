// Even-odd identities of the Wilson operator, exercised through the Grid-API facade.
//
// Each check is an algebraic property the red-black decomposition must satisfy for any gauge
// field and any vectors (the same properties Grid's tests/core/Test_wilson_even_odd.cc asserts):
//   split      the two half-lattice hops Meooe(x_odd) and Meooe(x_even) rebuild Dhop(x);
//   diagonal   M x = Mooee x + Dhop x on the full lattice;
//   adjoint    <u, Meooe v> = conj <v, MeooeDag u> for u, v on opposite checkerboards;
//   inverse    MooeeInv Mooee = 1 and MooeeInvDag MooeeDag = 1 on both checkerboards;
//   schur      A = MpcDagMpc is Hermitian (<u, A v> = conj <v, A u>) and positive (<v, A v> > 0).
#include <Grid/Grid.h>

#include <cstdio>
#include <functional>
#include <string>
#include <vector>

using namespace Grid;

namespace {

struct Check {
  std::string name;
  double value, bound;
};

std::vector<Check> results;

void record(const std::string &name, double value, double bound) { results.push_back({name, value, bound}); }

double rel(const LatticeFermionD &a, const LatticeFermionD &b) { return std::sqrt(norm2(a - b) / norm2(b)); }

double rel(ComplexD a, ComplexD b) { return std::abs(a - b) / std::abs(b); }

}  // namespace

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian *full = SpaceTimeGrid::makeFourDimGrid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridRedBlackCartesian *half = SpaceTimeGrid::makeFourDimRedBlackGrid(full);

  GridParallelRNG rng(full);
  rng.SeedFixedIntegers(std::vector<int>({7, 301, 4242, 15}));
  LatticeGaugeFieldD links(full);
  SU<Nc>::HotConfiguration(rng, links);
  WilsonFermionD op(links, *full, *half, 0.27);

  auto fresh = [&](GridBase *g) {
    LatticeFermionD v(g);
    gaussian(rng, v);
    return v;
  };
  auto parts = [&](const LatticeFermionD &x, LatticeFermionD &ev, LatticeFermionD &od) {
    pickCheckerboard(Even, ev, x);
    pickCheckerboard(Odd, od, x);
  };

  const LatticeFermionD x = fresh(full), u = fresh(full), v = fresh(full);
  LatticeFermionD xe(half), xo(half), ue(half), uo(half), ve(half), vo(half);
  parts(x, xe, xo);
  parts(u, ue, uo);
  parts(v, ve, vo);

  // split: the even part of Dhop x comes from the odd part of x and vice versa
  {
    LatticeFermionD whole(full), rebuilt(full), to_even(half), to_odd(half);
    op.Dhop(x, whole, DaggerNo);
    op.Meooe(xo, to_even);
    op.Meooe(xe, to_odd);
    setCheckerboard(rebuilt, to_even);
    setCheckerboard(rebuilt, to_odd);
    record("split: |Meooe(x_o) + Meooe(x_e) - Dhop x| / |Dhop x|", rel(rebuilt, whole), 1e-14);
    op.Dhop(x, whole, DaggerYes);
    op.MeooeDag(xo, to_even);
    op.MeooeDag(xe, to_odd);
    setCheckerboard(rebuilt, to_even);
    setCheckerboard(rebuilt, to_odd);
    record("split (dagger)", rel(rebuilt, whole), 1e-14);
  }

  // diagonal: M = Mooee + Dhop
  {
    LatticeFermionD mx(full), hop(full), sum(full), diag_e(half), diag_o(half), diag(full);
    op.M(x, mx);
    op.Dhop(x, hop, DaggerNo);
    op.Mooee(xe, diag_e);
    op.Mooee(xo, diag_o);
    setCheckerboard(diag, diag_e);
    setCheckerboard(diag, diag_o);
    sum = diag + hop;
    record("diagonal: |Mooee x + Dhop x - M x| / |M x|", rel(sum, mx), 1e-14);
  }

  // adjoint: <u_a, Meooe v_b> = conj <v_b, MeooeDag u_a>, a != b
  {
    LatticeFermionD dv_e(half), dv_o(half), du_e(half), du_o(half);
    op.Meooe(vo, dv_e);
    op.Meooe(ve, dv_o);
    op.MeooeDag(ue, du_o);
    op.MeooeDag(uo, du_e);
    record("adjoint, even target", rel(innerProduct(ue, dv_e), std::conj(innerProduct(vo, du_o))), 1e-13);
    record("adjoint, odd target", rel(innerProduct(uo, dv_o), std::conj(innerProduct(ve, du_e))), 1e-13);
  }

  // inverse: the site-diagonal block and its inverse, with and without dagger
  {
    const char *cb_name[2] = {"even", "odd"};
    LatticeFermionD *part[2] = {&ue, &uo};
    for (int cb = 0; cb < 2; cb++) {
      LatticeFermionD y(half), back(half);
      op.Mooee(*part[cb], y);
      op.MooeeInv(y, back);
      record(std::string("inverse: MooeeInv Mooee on ") + cb_name[cb], rel(back, *part[cb]), 1e-14);
      op.MooeeDag(*part[cb], y);
      op.MooeeInvDag(y, back);
      record(std::string("inverse: MooeeInvDag MooeeDag on ") + cb_name[cb], rel(back, *part[cb]), 1e-14);
    }
  }

  // schur: the normal operator of the Schur complement
  {
    SchurDiagMooeeOperator<WilsonFermionD, LatticeFermionD> schur(op);
    LatticeFermionD *us[2] = {&ue, &uo}, *vs[2] = {&ve, &vo};
    for (int cb = 0; cb < 2; cb++) {
      LatticeFermionD au(half), av(half);
      schur.MpcDagMpc(*us[cb], au);
      schur.MpcDagMpc(*vs[cb], av);
      const ComplexD uav = innerProduct(*us[cb], av), vau = innerProduct(*vs[cb], au);
      record(std::string("schur: Hermitian on ") + (cb ? "odd" : "even"), rel(uav, std::conj(vau)), 1e-13);
      const ComplexD vav = innerProduct(*vs[cb], av);
      record(std::string("schur: Im<v, A v> / Re<v, A v> on ") + (cb ? "odd" : "even"),
             std::abs(vav.imag()) / std::abs(vav.real()), 1e-13);
      record(std::string("schur: Re<v, A v> > 0 on ") + (cb ? "odd" : "even"), vav.real() > 0 ? 0.0 : 1.0, 0.5);
    }
  }

  int failed = 0;
  for (const Check &c : results) {
    const bool ok = c.value < c.bound;
    failed += ok ? 0 : 1;
    std::printf("%-60s %.3e (bound %.0e) %s\n", c.name.c_str(), c.value, c.bound, ok ? "ok" : "FAIL");
  }
  Grid_finalize();
  std::printf("%s\n", failed ? "FAILED" : "PASSED");
  return failed ? 1 : 0;
}
