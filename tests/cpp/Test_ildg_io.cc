// ILDG gauge files through the Grid-API facade (Grid's tests/IO/Test_ildg_io.cc writes a hot
// configuration with IldgWriter and reads it back with IldgReader): the field read back equals the
// one written, the header carries the LFN and the plaquette, and the plaquette agrees with
// WilsonLoops on the field.
#include <unistd.h>

#include <cstdio>

#include <Grid/Grid.h>

using namespace Grid;

int main(int argc, char **argv) {
  Grid_init(&argc, &argv);
  GridCartesian grid(GridDefaultLatt(), GridDefaultSimd(Nd, vComplex::Nsimd()), GridDefaultMpi());
  GridParallelRNG rng(&grid);
  rng.SeedFixedIntegers(std::vector<int>({3, 1, 4, 1}));
  LatticeGaugeField written(&grid), read_back(&grid);
  SU<Nc>::HotConfiguration(rng, written);
  const std::string dir = std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp";
  const std::string file = dir + "/ckpoint_ildg." + std::to_string((long)getpid()) + ".4000";

  IldgWriter writer(grid.IsBoss());
  writer.open(file);
  writer.writeConfiguration(written, 4000, "lfn://gx/test/4000", "hot configuration");
  writer.close();
  FieldMetaData header;
  IldgReader reader;
  reader.open(file);
  reader.readConfiguration(read_back, header);
  reader.close();
  std::remove(file.c_str());

  double diff = 0;
  for (size_t i = 0; i < written.data.size(); i++) diff += std::norm(written.data[i] - read_back.data[i]);
  const RealD plaq = WilsonLoops<PeriodicGimplR>::avgPlaquette(read_back);
  const bool ok = diff == 0 && header.ildg_lfn == "lfn://gx/test/4000" && std::fabs(header.plaquette - plaq) < 1e-12;
  std::printf("diff %.3e lfn %s plaquette %.12f (header %.12f)\n", diff, header.ildg_lfn.c_str(), plaq,
              header.plaquette);
  Grid_finalize();
  std::printf("%s\n", ok ? "PASSED" : "FAILED");
  return ok ? 0 : 1;
}
