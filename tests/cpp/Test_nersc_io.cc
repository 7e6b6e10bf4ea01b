// Mirrors Grid's tests/IO/Test_nersc_io.cc (gauge part; the RNG-state files and the sliced and
// coarsened plaquette sums are not part of this library): a hot configuration written as a NERSC
// file (three rows, then the two-row clone) reads back unchanged, and its plaquette and link trace
// agree with WilsonLoops on the field.
#include <unistd.h>

#include <cstdio>

#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian Fine(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridParallelRNG pRNGa(&Fine);
  pRNGa.SeedFixedIntegers(std::vector<int>({45, 12, 81, 9}));
  LatticeGaugeField Umu(&Fine), Umu_saved(&Fine);
  SU<Nc>::HotConfiguration(pRNGa, Umu);
  const std::string dir = std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp";
  const std::string tag = std::to_string((long)getpid());
  const std::string file = dir + "/ckpoint_lat." + tag + ".4000", clone2x3 = dir + "/ckpoint_clone2x3." + tag + ".4000";

  FieldMetaData header;
  NerscIO::writeConfiguration(Umu, file, 0, 0);
  Umu_saved = Umu;
  NerscIO::readConfiguration(Umu, header, file);
  double diff = 0;
  for (size_t i = 0; i < Umu.data.size(); i++) diff += std::norm(Umu.data[i] - Umu_saved.data[i]);
  const RealD plaq = WilsonLoops<PeriodicGimplR>::avgPlaquette(Umu), trace = WilsonLoops<PeriodicGimplR>::linkTrace(Umu);
  std::cout << GridLogMessage << "norm2 Gauge Diff = " << diff << "; calculated plaquettes " << plaq << " (header "
            << header.plaquette << "), link trace " << trace << " (header " << header.link_trace << ")" << std::endl;
  int fail = diff != 0.0 || !(std::fabs(plaq - header.plaquette) < 1e-9) || !(std::fabs(trace - header.link_trace) < 1e-9) ||
             header.data_type != "4D_SU3_GAUGE_3x3" || header.floating_point != "IEEE64BIG";

  NerscIO::writeConfiguration(Umu_saved, clone2x3, 1, 0);
  FieldMetaData h2;
  NerscIO::readConfiguration(Umu, h2, clone2x3);
  double d2 = 0, n2 = 0;
  for (size_t i = 0; i < Umu.data.size(); i++) {
    d2 += std::norm(Umu.data[i] - Umu_saved.data[i]);
    n2 += std::norm(Umu_saved.data[i]);
  }
  std::cout << GridLogMessage << "two-row clone: relative difference " << std::sqrt(d2 / n2) << std::endl;
  fail |= !(std::sqrt(d2 / n2) < 1e-14) || h2.data_type != "4D_SU3_GAUGE";
  std::remove(file.c_str());
  std::remove(clone2x3.c_str());
  Grid_finalize();
  std::cout << (fail ? "FAILED" : "PASSED") << std::endl;
  return fail;
}
