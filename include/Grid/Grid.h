// Grid/Grid.h — C++ facade with Grid's operator and solver API over libgx (include/gx.h).
//
// A Grid user's code for the even-odd Wilson / Mobius solve path compiles against this header
// unchanged in form: the same class names, constructors, method names, argument meaning,
// checkerboard tags and assert-on-error behaviour, with every operator application, BLAS1 call
// and solve executed by the HIP library on one MI355X.  Fermion fields live in device memory;
// gauge fields are host arrays handed to the operators (ImportGauge), as Grid's constructors take
// them.  Single process, single GPU: the MPI layout arguments are accepted and ignored.
//
// Reference interfaces (paths relative to the Grid source tree):
//   Grid_init / GridDefaultLatt           Grid/util/Init.cc:196-547
//   GridCartesian / GridRedBlackCartesian  Grid/cartesian/Cartesian_full.h, Cartesian_red_black.h:56-304
//   SpaceTimeGrid::make*Grid              Grid/qcd/utils/SpaceTimeGrid.cc:36-113
//   Lattice BLAS1 / reductions            Grid/lattice/Lattice_arith.h:231-265, Lattice_reduction.h:202-410
//   pick/setCheckerboard, precisionChange Grid/lattice/Lattice_transfer.h:50-80,1312-1492
//   CheckerBoardedSparseMatrixBase        Grid/algorithms/SparseMatrix.h:37-82
//   WilsonFermion                         Grid/qcd/action/fermion/WilsonFermion.h:139-142
//   MobiusFermion                         Grid/qcd/action/fermion/MobiusFermion.h:45-71
//   LinearOperatorBase / Schur operators  Grid/algorithms/LinearOperator.h:44-56,286-349
//   ConjugateGradient                     Grid/algorithms/iterative/ConjugateGradient.h:43-257
//   MixedPrecisionConjugateGradient       Grid/algorithms/iterative/ConjugateGradientMixedPrec.h:33-166
//   SchurRedBlackDiagMooeeSolve           Grid/algorithms/iterative/SchurRedBlack.h:236-436
#ifndef GRID_AMD_GRID_H
#define GRID_AMD_GRID_H

#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "../gx.h"

namespace Grid {

typedef double RealD;
typedef float RealF;
typedef RealD Real;
typedef std::complex<double> ComplexD;
typedef ComplexD Complex;
typedef int Integer;

static const int Nd = 4;
static const int Nc = 3;
static const int Ns = 4;
static const int Even = 0;
static const int Odd = 1;
static const int DaggerNo = 0;
static const int DaggerYes = 1;

#define GridLogMessage "Grid : Message : "
#define GridLogIterative "Grid : Iterative : "
#define GridLogPerformance "Grid : Performance : "
#define GridLogDebug "Grid : Debug : "

class Coordinate : public std::vector<int> {
 public:
  Coordinate() = default;
  Coordinate(std::initializer_list<int> l) : std::vector<int>(l) {}
  explicit Coordinate(size_t n, int v = 0) : std::vector<int>(n, v) {}
  Coordinate(const std::vector<int> &v) : std::vector<int>(v) {}
};

struct Zero {};

namespace detail {
inline gx_ctx &context() {
  static gx_ctx ctx = nullptr;
  return ctx;
}
inline Coordinate &default_latt() {
  static Coordinate latt({8, 8, 8, 8});
  return latt;
}
// Grid reports errors by assert; so does the facade, with the library's message first.
inline void check(int rc, const char *what) {
  if (rc != GX_OK) {
    std::cerr << "Grid (gx) error in " << what << ": " << gx_last_error() << " (status " << rc << ")" << std::endl;
    assert(rc == GX_OK);
    std::abort();
  }
}
// The device mixed-precision solvers run the sloppy solves with the FP64 operator's own links in
// lower precision, so they apply only when the sloppy operator is the same action.
inline bool same_action(gx_op d, gx_op f) {
  if (!d || !f) return false;
  int same = 0;
  check(gx_op_same_action(d, f, &same), "same_action");
  return same != 0;
}
// the libgx operator behind a matrix type, when it has one
template <class M>
auto handle_of(const M &m, int) -> decltype(m.handle()) {
  return m.handle();
}
template <class M>
gx_op handle_of(const M &, long) {
  return nullptr;
}
// small dense complex matrices of the block solvers' host algebra (Eigen's role in Grid), row-major
struct DenseC {
  int n = 0;
  std::vector<ComplexD> a;
  explicit DenseC(int n_ = 0, bool ident = false) : n(n_), a((size_t)n_ * n_) {
    for (int i = 0; ident && i < n; i++) a[(size_t)i * n + i] = 1.0;
  }
  ComplexD &operator()(int i, int j) { return a[(size_t)i * n + j]; }
  ComplexD operator()(int i, int j) const { return a[(size_t)i * n + j]; }
  DenseC adj() const {
    DenseC z(n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) z(i, j) = std::conj((*this)(j, i));
    return z;
  }
  DenseC operator*(const DenseC &y) const {
    DenseC z(n);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) z(i, j) += (*this)(i, k) * y(k, j);
    return z;
  }
  // Gauss-Jordan inverse with partial pivoting (asserts on a singular matrix, as Eigen's would fail)
  DenseC inverse() const {
    DenseC w = *this, out(n, true);
    for (int c = 0; c < n; c++) {
      int piv = c;
      for (int r = c + 1; r < n; r++)
        if (std::abs(w(r, c)) > std::abs(w(piv, c))) piv = r;
      assert(std::abs(w(piv, c)) > 0.0 && "singular block matrix");
      for (int j = 0; j < n && piv != c; j++) {
        std::swap(w(piv, j), w(c, j));
        std::swap(out(piv, j), out(c, j));
      }
      const ComplexD inv = 1.0 / w(c, c);
      for (int j = 0; j < n; j++) {
        w(c, j) *= inv;
        out(c, j) *= inv;
      }
      for (int r = 0; r < n; r++)
        if (r != c) {
          const ComplexD f = w(r, c);
          for (int j = 0; j < n; j++) {
            w(r, j) -= f * w(c, j);
            out(r, j) -= f * out(c, j);
          }
        }
    }
    return out;
  }
  // lower Cholesky factor L of a Hermitian positive-definite matrix (Eigen's llt().matrixL())
  DenseC cholesky() const {
    DenseC L(n);
    for (int j = 0; j < n; j++) {
      double d = (*this)(j, j).real();
      for (int k = 0; k < j; k++) d -= std::norm(L(j, k));
      assert(d > 0.0 && "block matrix not positive definite");
      L(j, j) = std::sqrt(d);
      for (int i = j + 1; i < n; i++) {
        ComplexD v = (*this)(i, j);
        for (int k = 0; k < j; k++) v -= L(i, k) * std::conj(L(j, k));
        L(i, j) = v / L(j, j);
      }
    }
    return L;
  }
};
inline uint64_t mix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}
}  // namespace detail

// Grid_init: parses --grid X.Y.Z.T and --device N; creates the HIP context.
inline void Grid_init(int *argc, char ***argv) {
  int device = 0;
  for (int i = 1; argc && i < *argc; i++) {
    std::string a = (*argv)[i];
    if (a == "--grid" && i + 1 < *argc) {
      Coordinate c;
      std::stringstream ss((*argv)[i + 1]);
      std::string tok;
      while (std::getline(ss, tok, '.')) c.push_back(std::atoi(tok.c_str()));
      if (c.size() == 4) detail::default_latt() = c;
    } else if (a == "--device" && i + 1 < *argc) {
      device = std::atoi((*argv)[i + 1]);
    }
  }
  if (!detail::context()) detail::check(gx_ctx_create(device, &detail::context()), "Grid_init");
}
// Grid_finalize: waits for the device.  The context stays alive until the process exits, because
// Grid programs call Grid_finalize() while their fields and operators are still in scope.
inline void Grid_finalize() {
  if (detail::context()) gx_ctx_sync(detail::context());
}
inline Coordinate GridDefaultLatt() { return detail::default_latt(); }
inline Coordinate GridDefaultSimd(int nd, int) { return Coordinate(nd, 1); }
inline Coordinate GridDefaultMpi() { return Coordinate(4, 1); }
struct vComplex {
  static int Nsimd() { return 1; }
};
typedef vComplex vComplexD;
typedef vComplex vComplexF;
struct GridThread {
  static int GetThreads() { return 1; }
};

// ---------------- grids ----------------
class GridBase {
 public:
  bool IsBoss() const { return true; }  // one process per GPU: this process does the IO
  int dims[4] = {0, 0, 0, 0};
  int Ls = 1;
  bool checkerboarded = false;
  virtual ~GridBase() = default;
  int isCheckerBoarded() const { return checkerboarded ? 1 : 0; }
  // Grid's 5D grids put s first: (Ls, x, y, z, t)
  Coordinate GlobalDimensions() const {
    Coordinate c;
    if (Ls > 1) c.push_back(Ls);
    for (int d = 0; d < 4; d++) c.push_back(dims[d]);
    return c;
  }
  int Nd() const { return Ls > 1 ? 5 : 4; }
  int64_t gSites() const { return (int64_t)dims[0] * dims[1] * dims[2] * dims[3] * Ls / (checkerboarded ? 2 : 1); }
  int64_t oSites() const { return gSites(); }
  int64_t lSites() const { return gSites(); }
};

class GridCartesian : public GridBase {
 public:
  GridCartesian(const Coordinate &latt, const Coordinate & = Coordinate(), const Coordinate & = Coordinate()) {
    assert(latt.size() == 4);
    for (int d = 0; d < 4; d++) dims[d] = latt[d];
  }
  GridCartesian(int ls, const GridCartesian *four) {
    for (int d = 0; d < 4; d++) dims[d] = four->dims[d];
    Ls = ls;
  }
};

class GridRedBlackCartesian : public GridBase {
 public:
  explicit GridRedBlackCartesian(const GridBase *base) {
    for (int d = 0; d < 4; d++) dims[d] = base->dims[d];
    Ls = base->Ls;
    checkerboarded = true;
    assert(dims[0] % 2 == 0);
  }
};

namespace SpaceTimeGrid {
inline GridCartesian *makeFourDimGrid(const Coordinate &latt, const Coordinate &simd, const Coordinate &mpi) {
  return new GridCartesian(latt, simd, mpi);
}
inline GridRedBlackCartesian *makeFourDimRedBlackGrid(const GridCartesian *g) { return new GridRedBlackCartesian(g); }
inline GridCartesian *makeFiveDimGrid(int Ls, const GridCartesian *g) { return new GridCartesian(Ls, g); }
inline GridRedBlackCartesian *makeFiveDimRedBlackGrid(int Ls, const GridCartesian *g) {
  GridCartesian five(Ls, g);
  return new GridRedBlackCartesian(&five);
}
}  // namespace SpaceTimeGrid

// ---------------- random numbers ----------------
class GridParallelRNG {
 public:
  uint64_t seed = 0x5EED;
  uint64_t draws = 0;
  explicit GridParallelRNG(const GridBase *) {}
  void SeedFixedIntegers(const std::vector<int> &seeds) {
    uint64_t s = 0x123456789ABCDEFULL;
    for (int v : seeds) s = detail::mix(s ^ (uint64_t)(uint32_t)v);
    seed = s;
    draws = 0;
  }
  uint64_t next() { return detail::mix(seed + 0x1000193ULL * (++draws)); }
};
typedef GridParallelRNG GridSerialRNG;

// ---------------- fermion fields (device resident) ----------------
template <int PREC>
class FermionLattice {
 public:
  static constexpr int Prec = PREC;
  GridBase *_grid = nullptr;
  gx_field f = nullptr;

  explicit FermionLattice(GridBase *g) : _grid(g) {
    detail::check(gx_field_create(detail::context(), g->dims, g->Ls, PREC, g->checkerboarded ? GX_EVEN : GX_FULL, &f),
                  "Lattice(grid)");
  }
  FermionLattice(const FermionLattice &o) : FermionLattice(o._grid) { *this = o; }
  FermionLattice(FermionLattice &&o) noexcept : _grid(o._grid), f(o.f) { o.f = nullptr; }
  ~FermionLattice() {
    if (f) gx_field_destroy(f);
  }
  FermionLattice &operator=(const FermionLattice &o) {
    if (this != &o) detail::check(gx_field_copy(f, o.f), "Lattice copy");
    return *this;
  }
  FermionLattice &operator=(FermionLattice &&o) noexcept {
    std::swap(f, o.f);
    _grid = o._grid;
    return *this;
  }
  FermionLattice &operator=(const Zero &) {
    detail::check(gx_field_zero(f), "Lattice = Zero()");
    return *this;
  }
  GridBase *Grid() const { return _grid; }
  // Grid: Lattice::Checkerboard() (read and assign)
  struct CbRef {
    gx_field f;
    operator int() const {
      int cb;
      gx_field_checkerboard(f, &cb);
      return cb < 0 ? 0 : cb;
    }
    CbRef &operator=(int cb) {
      int cur;
      gx_field_checkerboard(f, &cur);
      if (cur >= 0) gx_field_set_checkerboard_tag(f, cb);
      return *this;
    }
  };
  CbRef Checkerboard() { return CbRef{f}; }
  int Checkerboard() const {
    int cb;
    gx_field_checkerboard(f, &cb);
    return cb < 0 ? 0 : cb;
  }
  // host transfer in Grid's lexicographic layout (s innermost), complex doubles
  size_t hostSize() const { return (size_t)_grid->gSites() * 12; }
  void importHost(const ComplexD *p) { detail::check(gx_field_import(f, reinterpret_cast<const double *>(p)), "import"); }
  void exportHost(ComplexD *p) const { detail::check(gx_field_export(f, reinterpret_cast<double *>(p)), "export"); }
};

typedef FermionLattice<GX_F64> LatticeFermionD;
typedef FermionLattice<GX_F32> LatticeFermionF;
typedef LatticeFermionD LatticeFermion;

template <int P>
inline RealD norm2(const FermionLattice<P> &x) {
  double n;
  detail::check(gx_norm2(x.f, &n), "norm2");
  return n;
}
template <int P>
inline ComplexD innerProduct(const FermionLattice<P> &x, const FermionLattice<P> &y) {
  double re, im;
  detail::check(gx_inner_product(x.f, y.f, &re, &im), "innerProduct");
  return ComplexD(re, im);
}
template <int P>
inline void innerProductNorm(ComplexD &ip, RealD &nrm, const FermionLattice<P> &x, const FermionLattice<P> &y) {
  double re, im;
  detail::check(gx_inner_product_norm(x.f, y.f, &re, &im, &nrm), "innerProductNorm");
  ip = ComplexD(re, im);
}
template <int P>
inline void axpy(FermionLattice<P> &z, RealD a, const FermionLattice<P> &x, const FermionLattice<P> &y) {
  detail::check(gx_axpy(z.f, a, x.f, y.f), "axpy");
}
template <int P>
inline void axpby(FermionLattice<P> &z, RealD a, RealD b, const FermionLattice<P> &x, const FermionLattice<P> &y) {
  detail::check(gx_axpby(z.f, a, b, x.f, y.f), "axpby");
}
template <int P>
inline RealD axpy_norm(FermionLattice<P> &z, RealD a, const FermionLattice<P> &x, const FermionLattice<P> &y) {
  double n;
  detail::check(gx_axpy_norm(z.f, a, x.f, y.f, &n), "axpy_norm");
  return n;
}
template <int P>
inline FermionLattice<P> operator-(const FermionLattice<P> &a, const FermionLattice<P> &b) {
  FermionLattice<P> r(a.Grid());
  axpby(r, 1.0, -1.0, a, b);
  return r;
}
template <int P>
inline FermionLattice<P> operator+(const FermionLattice<P> &a, const FermionLattice<P> &b) {
  FermionLattice<P> r(a.Grid());
  axpby(r, 1.0, 1.0, a, b);
  return r;
}
template <int P>
inline FermionLattice<P> operator*(RealD s, const FermionLattice<P> &a) {
  FermionLattice<P> r(a.Grid());
  detail::check(gx_scale(r.f, s, a.f), "scale");
  return r;
}
template <int P>
inline void pickCheckerboard(int cb, FermionLattice<P> &half, const FermionLattice<P> &full) {
  detail::check(gx_pick_checkerboard(cb, half.f, full.f), "pickCheckerboard");
}
template <int P>
inline void setCheckerboard(FermionLattice<P> &full, const FermionLattice<P> &half) {
  detail::check(gx_set_checkerboard(full.f, half.f), "setCheckerboard");
}
template <int P, int Q>
inline void precisionChange(FermionLattice<P> &out, const FermionLattice<Q> &in) {
  detail::check(gx_precision_change(out.f, in.f), "precisionChange");
}
template <int P>
inline void gaussian(GridParallelRNG &rng, FermionLattice<P> &x) {
  GridBase *g = x.Grid();
  GridCartesian full(Coordinate({g->dims[0], g->dims[1], g->dims[2], g->dims[3]}));
  std::vector<ComplexD> host((size_t)g->dims[0] * g->dims[1] * g->dims[2] * g->dims[3] * g->Ls * 12);
  detail::check(gx_gaussian_fermion(g->dims, g->Ls, rng.next(), reinterpret_cast<double *>(host.data())), "gaussian");
  if (!g->checkerboarded) {
    x.importHost(host.data());
    return;
  }
  // a half field takes the sites of its checkerboard (x halved, Grid red-black order)
  const int cb = x.Checkerboard();
  std::vector<ComplexD> half(host.size() / 2);
  const int lx = g->dims[0], ly = g->dims[1], lz = g->dims[2], lt = g->dims[3], Ls = g->Ls;
  for (int t = 0; t < lt; t++)
    for (int z = 0; z < lz; z++)
      for (int y = 0; y < ly; y++)
        for (int xx = 0; xx < lx; xx++) {
          if (((xx + y + z + t) & 1) != cb) continue;
          const int64_t s4 = xx + (int64_t)lx * (y + (int64_t)ly * (z + (int64_t)lz * t));
          const int64_t h = (xx >> 1) + (int64_t)(lx / 2) * (y + (int64_t)ly * (z + (int64_t)lz * t));
          std::memcpy(&half[h * Ls * 12], &host[s4 * Ls * 12], sizeof(ComplexD) * Ls * 12);
        }
  x.importHost(half.data());
}
template <int P>
inline void random(GridParallelRNG &rng, FermionLattice<P> &x) {
  gaussian(rng, x);
}

// ---------------- gauge fields (host resident, Grid lexicographic [site][mu][a][b]) ----------------
class LatticeGaugeFieldD {
 public:
  GridBase *_grid;
  std::vector<ComplexD> data;
  explicit LatticeGaugeFieldD(GridBase *g) : _grid(g), data((size_t)g->dims[0] * g->dims[1] * g->dims[2] * g->dims[3] * 36) {}
  GridBase *Grid() const { return _grid; }
  LatticeGaugeFieldD &operator=(const Zero &) {
    std::fill(data.begin(), data.end(), ComplexD(0));
    return *this;
  }
  const double *raw() const { return reinterpret_cast<const double *>(data.data()); }
  double *raw() { return reinterpret_cast<double *>(data.data()); }
};
typedef LatticeGaugeFieldD LatticeGaugeFieldF;
typedef LatticeGaugeFieldD LatticeGaugeField;
inline void precisionChange(LatticeGaugeFieldD &out, const LatticeGaugeFieldD &in) { out.data = in.data; }

// Grid: WilsonLoops<Gimpl>::avgPlaquette / linkTrace (Grid/qcd/utils/WilsonLoops.h:121-126,203-219)
template <class Gimpl = void>
struct WilsonLoops {
  static RealD avgPlaquette(const LatticeGaugeFieldD &U) {
    double p = 0;
    detail::check(gx_gauge_statistics(detail::context(), U.Grid()->dims, U.raw(), &p, nullptr), "avgPlaquette");
    return p;
  }
  static RealD linkTrace(const LatticeGaugeFieldD &U) {
    double t = 0;
    detail::check(gx_gauge_statistics(detail::context(), U.Grid()->dims, U.raw(), nullptr, &t), "linkTrace");
    return t;
  }
};
typedef void PeriodicGimplR;

// Grid: FieldMetaData (Grid/parallelIO/MetaData.h:62-98), the NERSC header fields a reader returns
struct FieldMetaData {
  std::vector<int> dimension = std::vector<int>(4);
  std::string data_type, floating_point, ensemble_id, ensemble_label, ildg_lfn;
  RealD link_trace = 0, plaquette = 0;
  uint32_t checksum = 0, scidac_checksuma = 0, scidac_checksumb = 0;
  int sequence_number = 0;
};

// Grid: NerscIO::readConfiguration / writeConfiguration (Grid/parallelIO/NerscIO.h:148-290).  A
// checksum, plaquette or link-trace mismatch asserts, as in Grid.
// Grid: IldgWriter(isboss).open(file); writeConfiguration(Umu, sequence, LFN, description); close()
// (IldgIO.h:611-690).  Writes IEEE64BIG as Grid does; the sequence number lives in Grid's private
// grid-format record, which this writer does not produce.
class IldgWriter {
 public:
  explicit IldgWriter(bool = true) {}
  void open(const std::string &file) { filename = file; }
  void writeConfiguration(LatticeGaugeFieldD &Umu, int sequence, const std::string &LFN, const std::string &description) {
    (void)sequence;
    detail::check(gx_ildg_write(detail::context(), filename.c_str(), Umu.Grid()->dims, Umu.raw(), 64, LFN.c_str(),
                                description.c_str()),
                  "IldgWriter::writeConfiguration");
  }
  void close() {}

 private:
  std::string filename;
};
// Grid: IldgReader.open(file); readConfiguration(Umu, FieldMetaData&); close() (IldgIO.h:698-936)
class IldgReader {
 public:
  void open(const std::string &file) { filename = file; }
  void readConfiguration(LatticeGaugeFieldD &Umu, FieldMetaData &header) {
    gx_ildg_info info{};
    const int rc = gx_ildg_read(detail::context(), filename.c_str(), Umu.Grid()->dims, Umu.raw(), &info);
    header.dimension.assign(info.dims, info.dims + 4);
    header.data_type = "4D_SU3_GAUGE_3x3";
    header.floating_point = info.precision == 32 ? "IEEE32BIG" : "IEEE64BIG";
    header.ildg_lfn = info.lfn;
    header.plaquette = info.has_plaquette ? info.plaquette_file : 0.0;
    header.link_trace = info.has_plaquette ? info.link_trace_file : 0.0;
    header.scidac_checksuma = info.scidac_checksuma_file;
    header.scidac_checksumb = info.scidac_checksumb_file;
    detail::check(rc, "IldgReader::readConfiguration");
    std::cout << GridLogMessage << "ILDG configuration " << filename << ": SciDAC checksums " << std::hex
              << info.scidac_checksuma << " " << info.scidac_checksumb << std::dec << " match; plaquette "
              << info.plaquette << " link trace " << info.link_trace << std::endl;
  }
  void close() {}

 private:
  std::string filename;
};

struct NerscIO {
  static void readConfiguration(LatticeGaugeFieldD &Umu, FieldMetaData &header, const std::string &file) {
    gx_nersc_info info{};
    const int rc = gx_nersc_read(detail::context(), file.c_str(), Umu.Grid()->dims, Umu.raw(), &info);
    header.dimension.assign(info.dims, info.dims + 4);
    header.data_type = info.two_row ? "4D_SU3_GAUGE" : "4D_SU3_GAUGE_3x3";
    header.floating_point = info.floating_point;
    header.ensemble_id = info.ensemble_id;
    header.ensemble_label = info.ensemble_label;
    header.link_trace = info.link_trace_header;
    header.plaquette = info.plaquette_header;
    header.checksum = info.checksum_header;
    header.sequence_number = info.sequence_number;
    detail::check(rc, "NerscIO::readConfiguration");
    std::cout << GridLogMessage << "NERSC Configuration " << file << " checksum " << std::hex << info.checksum_data
              << std::dec << " plaquette " << info.plaquette << " link_trace " << info.link_trace
              << ": plaquette, link trace, and checksum agree" << std::endl;
  }
  static void writeConfiguration(LatticeGaugeFieldD &Umu, const std::string &file, int two_row = 0, int bits32 = 1,
                                 const std::string &ens_label = "DWF", const std::string &ens_id = "UKQCD",
                                 unsigned int sequence_number = 1) {
    (void)bits32;  // Grid writes IEEE64BIG whatever bits32 says
    detail::check(gx_nersc_write(detail::context(), file.c_str(), Umu.Grid()->dims, Umu.raw(), two_row,
                                 ens_label.c_str(), ens_id.c_str(), (int)sequence_number),
                  "NerscIO::writeConfiguration");
  }
};

template <int N>
struct SU {
  // Random SU(3) links exp(i eps H), H Hermitian traceless Gaussian; eps = 1 for a hot start.
  static void HotConfiguration(GridParallelRNG &rng, LatticeGaugeFieldD &U, double eps = 1.0) {
    detail::check(gx_random_gauge(U.Grid()->dims, rng.next(), eps, U.raw()), "HotConfiguration");
  }
  static void ColdConfiguration(GridParallelRNG &, LatticeGaugeFieldD &U) { ColdConfiguration(U); }
  static void ColdConfiguration(LatticeGaugeFieldD &U) {
    std::fill(U.data.begin(), U.data.end(), ComplexD(0));
    for (size_t l = 0; l < U.data.size() / 9; l++)
      for (int a = 0; a < 3; a++) U.data[l * 9 + a * 4] = 1.0;
  }
};
typedef SU<3> SU3;

// ---------------- operators ----------------
struct WilsonImplParams {
  std::vector<ComplexD> boundary_phases{1, 1, 1, 1};
};
typedef WilsonImplParams ImplParams;

template <class Field>
class SparseMatrixBase {
 public:
  virtual GridBase *Grid() = 0;
  virtual void M(const Field &in, Field &out) = 0;
  virtual void Mdag(const Field &in, Field &out) = 0;
  virtual void MdagM(const Field &in, Field &out) {
    Field tmp(in.Grid());
    M(in, tmp);
    Mdag(tmp, out);
  }
  virtual ~SparseMatrixBase() = default;
};

template <class Field>
class CheckerBoardedSparseMatrixBase : public SparseMatrixBase<Field> {
 public:
  virtual GridBase *RedBlackGrid() = 0;
  virtual RealD Mass() { return 0.0; }
  virtual void Meooe(const Field &in, Field &out) = 0;
  virtual void Mooee(const Field &in, Field &out) = 0;
  virtual void MooeeInv(const Field &in, Field &out) = 0;
  virtual void MeooeDag(const Field &in, Field &out) = 0;
  virtual void MooeeDag(const Field &in, Field &out) = 0;
  virtual void MooeeInvDag(const Field &in, Field &out) = 0;
  // the underlying libgx operator (lets the solvers run device-resident)
  virtual gx_op handle() const = 0;
};

// Common body of WilsonFermion / MobiusFermion: a gx operator plus Grid's method names.
template <int PREC>
class FermionOperatorBase : public CheckerBoardedSparseMatrixBase<FermionLattice<PREC>> {
 public:
  typedef FermionLattice<PREC> FermionField;
  typedef LatticeGaugeFieldD GaugeField;
  GridBase *_FGrid, *_FrbGrid;
  gx_op op = nullptr;
  RealD mass;
  FermionOperatorBase(GridBase *fg, GridBase *frb, RealD m) : _FGrid(fg), _FrbGrid(frb), mass(m) {}
  ~FermionOperatorBase() override {
    if (op) gx_op_destroy(op);
  }
  FermionOperatorBase(const FermionOperatorBase &) = delete;
  FermionOperatorBase &operator=(const FermionOperatorBase &) = delete;
  gx_op handle() const override { return op; }
  GridBase *Grid() override { return _FGrid; }
  GridBase *RedBlackGrid() override { return _FrbGrid; }
  GridBase *FermionGrid() { return _FGrid; }
  GridBase *FermionRedBlackGrid() { return _FrbGrid; }
  RealD Mass() override { return mass; }
  void ImportGauge(const GaugeField &U) { detail::check(gx_op_import_gauge(op, U.raw()), "ImportGauge"); }
  void M(const FermionField &in, FermionField &out) override { detail::check(gx_M(op, in.f, out.f), "M"); }
  void Mdag(const FermionField &in, FermionField &out) override { detail::check(gx_Mdag(op, in.f, out.f), "Mdag"); }
  void Meooe(const FermionField &in, FermionField &out) override { detail::check(gx_Meooe(op, in.f, out.f), "Meooe"); }
  void MeooeDag(const FermionField &in, FermionField &out) override {
    detail::check(gx_MeooeDag(op, in.f, out.f), "MeooeDag");
  }
  void Mooee(const FermionField &in, FermionField &out) override { detail::check(gx_Mooee(op, in.f, out.f), "Mooee"); }
  void MooeeDag(const FermionField &in, FermionField &out) override {
    detail::check(gx_MooeeDag(op, in.f, out.f), "MooeeDag");
  }
  void MooeeInv(const FermionField &in, FermionField &out) override {
    detail::check(gx_MooeeInv(op, in.f, out.f), "MooeeInv");
  }
  void MooeeInvDag(const FermionField &in, FermionField &out) override {
    detail::check(gx_MooeeInvDag(op, in.f, out.f), "MooeeInvDag");
  }
  void Mdiag(const FermionField &in, FermionField &out) { Mooee(in, out); }
  void Dhop(const FermionField &in, FermionField &out, int dag) { detail::check(gx_Dhop(op, in.f, out.f, dag), "Dhop"); }
  void DhopEO(const FermionField &in, FermionField &out, int dag) {
    detail::check(gx_DhopEO(op, in.f, out.f, dag), "DhopEO");
  }
  void DhopOE(const FermionField &in, FermionField &out, int dag) {
    detail::check(gx_DhopOE(op, in.f, out.f, dag), "DhopOE");
  }
  // DhopDir / Mdir (FermionOperator.h:78,92): one leg of the hop, dir 0..3, disp +-1
  void DhopDir(const FermionField &in, FermionField &out, int dir, int disp) {
    out.Checkerboard() = in.Checkerboard();
    detail::check(gx_DhopDir(op, in.f, out.f, dir, disp), "DhopDir");
  }
  void Mdir(const FermionField &in, FermionField &out, int dir, int disp) { DhopDir(in, out, dir, disp); }
  // DhopDirAll / MdirAll (WilsonFermionImplementation.h:350-384): out[dir + 4 * (disp == -1)], 8 or 9 fields
  void DhopDirAll(const FermionField &in, std::vector<FermionField> &out) {
    assert(out.size() == 8 || out.size() == 9);
    for (int dir = 0; dir < Nd; dir++)
      for (int disp = -1; disp <= 1; disp += 2) DhopDir(in, out[dir + (disp == 1 ? 0 : 4)], dir, disp);
  }
  void MdirAll(const FermionField &in, std::vector<FermionField> &out) { DhopDirAll(in, out); }
  // the force terms (FermionOperator.h:81-89): mat is the host gauge field the derivative lands in
  void DhopDeriv(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) {
    detail::check(gx_DhopDeriv(op, mat.raw(), U.f, V.f, dag), "DhopDeriv");
  }
  void DhopDerivEO(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) {
    detail::check(gx_DhopDerivEO(op, mat.raw(), U.f, V.f, dag), "DhopDerivEO");
  }
  void DhopDerivOE(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) {
    detail::check(gx_DhopDerivOE(op, mat.raw(), U.f, V.f, dag), "DhopDerivOE");
  }
  void MDeriv(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) { DhopDeriv(mat, U, V, dag); }
  void MoeDeriv(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) { DhopDerivOE(mat, U, V, dag); }
  void MeoDeriv(GaugeField &mat, const FermionField &U, const FermionField &V, int dag) { DhopDerivEO(mat, U, V, dag); }

 protected:
  static const double *phases_of(const ImplParams &p, std::vector<double> &buf) {
    buf.resize(8);
    for (int mu = 0; mu < 4; mu++) {
      buf[2 * mu] = p.boundary_phases[mu].real();
      buf[2 * mu + 1] = p.boundary_phases[mu].imag();
    }
    return buf.data();
  }
};

// Grid: WilsonFermion<Impl>(GaugeField&, GridCartesian&, GridRedBlackCartesian&, RealD mass, params)
template <int PREC>
class WilsonFermion : public FermionOperatorBase<PREC> {
 public:
  WilsonFermion(const LatticeGaugeFieldD &U, GridCartesian &Fgrid, GridRedBlackCartesian &Hgrid, RealD mass,
                const ImplParams &p = ImplParams())
      : FermionOperatorBase<PREC>(&Fgrid, &Hgrid, mass) {
    std::vector<double> ph;
    detail::check(gx_wilson_create(detail::context(), Fgrid.dims, Fgrid.Ls, mass, U.raw(),
                                   FermionOperatorBase<PREC>::phases_of(p, ph), &this->op),
                  "WilsonFermion");
  }
};
typedef WilsonFermion<GX_F64> WilsonFermionD;
typedef WilsonFermion<GX_F32> WilsonFermionF;
typedef WilsonFermionD WilsonFermionR;

// Grid: MobiusFermion(U, FiveDimGrid, FiveDimRedBlackGrid, FourDimGrid, FourDimRedBlackGrid, mass, M5, b, c)
template <int PREC>
class MobiusFermion : public FermionOperatorBase<PREC> {
 public:
  RealD M5, _b, _c;
  MobiusFermion(const LatticeGaugeFieldD &U, GridCartesian &FiveDimGrid, GridRedBlackCartesian &FiveDimRedBlackGrid,
                GridCartesian &FourDimGrid, GridRedBlackCartesian &, RealD mass, RealD m5, RealD b, RealD c,
                const ImplParams &p = ImplParams())
      : FermionOperatorBase<PREC>(&FiveDimGrid, &FiveDimRedBlackGrid, mass), M5(m5), _b(b), _c(c) {
    assert(FiveDimGrid.Ls > 1);
    for (int d = 0; d < 4; d++) assert(FourDimGrid.dims[d] == FiveDimGrid.dims[d]);
    std::vector<double> ph;
    detail::check(gx_mobius_create(detail::context(), FiveDimGrid.dims, FiveDimGrid.Ls, mass, m5, b, c, U.raw(),
                                   FermionOperatorBase<PREC>::phases_of(p, ph), &this->op),
                  "MobiusFermion");
  }
  typedef FermionLattice<PREC> FermionField;
  void M5D(const FermionField &psi, FermionField &chi) { detail::check(gx_M5D(this->op, psi.f, chi.f), "M5D"); }
  void M5Ddag(const FermionField &psi, FermionField &chi) {
    detail::check(gx_M5Ddag(this->op, psi.f, chi.f), "M5Ddag");
  }
  void Meooe5D(const FermionField &psi, FermionField &chi) {
    detail::check(gx_Meooe5D(this->op, psi.f, chi.f), "Meooe5D");
  }
  void MeooeDag5D(const FermionField &psi, FermionField &chi) {
    detail::check(gx_MeooeDag5D(this->op, psi.f, chi.f), "MeooeDag5D");
  }
};
typedef MobiusFermion<GX_F64> MobiusFermionD;
typedef MobiusFermion<GX_F32> MobiusFermionF;
typedef MobiusFermionD MobiusFermionR;

// Grid: DomainWallFermion(U, FGrid, FrbGrid, UGrid, UrbGrid, mass, M5), DomainWallFermion.h:120-135:
// the Shamir kernel, i.e. the Cayley form with b = 1, c = 0.
template <int PREC>
class DomainWallFermion : public MobiusFermion<PREC> {
 public:
  DomainWallFermion(const LatticeGaugeFieldD &U, GridCartesian &FiveDimGrid, GridRedBlackCartesian &FiveDimRedBlackGrid,
                    GridCartesian &FourDimGrid, GridRedBlackCartesian &FourDimRedBlackGrid, RealD mass, RealD m5,
                    const ImplParams &p = ImplParams())
      : MobiusFermion<PREC>(U, FiveDimGrid, FiveDimRedBlackGrid, FourDimGrid, FourDimRedBlackGrid, mass, m5, 1.0, 0.0,
                            p) {}
};
typedef DomainWallFermion<GX_F64> DomainWallFermionD;
typedef DomainWallFermion<GX_F32> DomainWallFermionF;
typedef DomainWallFermionD DomainWallFermionR;

// ---------------- linear operators ----------------
template <class Field>
class LinearOperatorBase {
 public:
  virtual void Op(const Field &in, Field &out) = 0;
  virtual void AdjOp(const Field &in, Field &out) = 0;
  virtual void HermOpAndNorm(const Field &in, Field &out, RealD &n1, RealD &n2) = 0;
  virtual void HermOp(const Field &in, Field &out) = 0;
  // set when the operator is a Schur operator over a libgx action: solvers then run on device
  virtual gx_op schur_handle() const { return nullptr; }
  // set when the operator is Mdag M of a libgx action (full fields): block solvers run on device
  virtual gx_op mdagm_handle() const { return nullptr; }
  virtual ~LinearOperatorBase() = default;
};

template <class Field>
class SchurOperatorBase : public LinearOperatorBase<Field> {
 public:
  virtual void Mpc(const Field &in, Field &out) = 0;
  virtual void MpcDag(const Field &in, Field &out) = 0;
  virtual void MpcDagMpc(const Field &in, Field &out) {
    Field tmp(in.Grid());
    tmp.Checkerboard() = in.Checkerboard();
    Mpc(in, tmp);
    MpcDag(tmp, out);
  }
  void HermOpAndNorm(const Field &in, Field &out, RealD &n1, RealD &n2) override {
    MpcDagMpc(in, out);
    ComplexD dot = innerProduct(in, out);
    n1 = dot.real();
    n2 = norm2(out);
  }
  void HermOp(const Field &in, Field &out) override { MpcDagMpc(in, out); }
  void Op(const Field &in, Field &out) override { Mpc(in, out); }
  void AdjOp(const Field &in, Field &out) override { MpcDag(in, out); }
};

// Grid: SchurDiagMooeeOperator<Matrix,Field>(Matrix&), Mpc = Mooee - Meo MooeeInv Moe
template <class Matrix, class Field>
class SchurDiagMooeeOperator : public SchurOperatorBase<Field> {
 public:
  Matrix &_Mat;
  explicit SchurDiagMooeeOperator(Matrix &Mat) : _Mat(Mat) {}
  void Mpc(const Field &in, Field &out) override { detail::check(gx_Mpc(_Mat.handle(), in.f, out.f), "Mpc"); }
  void MpcDag(const Field &in, Field &out) override { detail::check(gx_MpcDag(_Mat.handle(), in.f, out.f), "MpcDag"); }
  void MpcDagMpc(const Field &in, Field &out) override {
    detail::check(gx_MpcDagMpc(_Mat.handle(), in.f, out.f), "MpcDagMpc");
  }
  void HermOpAndNorm(const Field &in, Field &out, RealD &n1, RealD &n2) override {
    detail::check(gx_HermOpAndNorm(_Mat.handle(), in.f, out.f, &n1, &n2), "HermOpAndNorm");
  }
  gx_op schur_handle() const override { return _Mat.handle(); }
};

// Grid: MdagMLinearOperator<Matrix,Field> (unpreconditioned normal operator)
template <class Matrix, class Field>
class MdagMLinearOperator : public LinearOperatorBase<Field> {
 public:
  Matrix &_Mat;
  explicit MdagMLinearOperator(Matrix &Mat) : _Mat(Mat) {}
  void Op(const Field &in, Field &out) override { _Mat.M(in, out); }
  void AdjOp(const Field &in, Field &out) override { _Mat.Mdag(in, out); }
  void HermOp(const Field &in, Field &out) override {
    Field tmp(in.Grid());
    _Mat.M(in, tmp);
    _Mat.Mdag(tmp, out);
  }
  void HermOpAndNorm(const Field &in, Field &out, RealD &n1, RealD &n2) override {
    HermOp(in, out);
    n1 = innerProduct(in, out).real();
    n2 = norm2(out);
  }
  gx_op mdagm_handle() const override { return detail::handle_of(_Mat, 0); }
};

template <class Field>
class OperatorFunction {
 public:
  virtual void operator()(LinearOperatorBase<Field> &Linop, const Field &in, Field &out) = 0;
  virtual ~OperatorFunction() = default;
};

// Grid: ConjugateGradient<Field>(tol, maxit, err_on_no_conv).  On a Schur operator over a libgx
// action the whole iteration runs on the device (gx_cg); on any other LinearOperatorBase the same
// Hestenes-Stiefel loop (ConjugateGradient.h:68-257) runs here over the facade's BLAS1.
template <class Field>
class ConjugateGradient : public OperatorFunction<Field> {
 public:
  RealD Tolerance;
  Integer MaxIterations;
  bool ErrorOnNoConverge;
  Integer IterationsToComplete = 0;
  RealD TrueResidual = 0;
  ConjugateGradient(RealD tol, Integer maxit, bool err_on_no_conv = true)
      : Tolerance(tol), MaxIterations(maxit), ErrorOnNoConverge(err_on_no_conv) {}

  void operator()(LinearOperatorBase<Field> &Linop, const Field &src, Field &psi) override {
    psi.Checkerboard() = src.Checkerboard();
    if (gx_op op = Linop.schur_handle()) {
      gx_solver_stats st{};
      const int rc = gx_cg(op, src.f, psi.f, Tolerance, MaxIterations, &st);
      IterationsToComplete = st.iterations;
      TrueResidual = st.true_residual;
      std::cout << GridLogMessage << "ConjugateGradient Converged on iteration " << st.iterations
                << "\tComputed residual " << st.computed_residual << "\tTrue residual " << st.true_residual
                << "\tTarget " << Tolerance << std::endl;
      if (rc == GX_ENOCONV) {
        std::cout << GridLogMessage << "ConjugateGradient did NOT converge " << MaxIterations << std::endl;
        if (ErrorOnNoConverge) assert(0);
        return;
      }
      detail::check(rc, "ConjugateGradient");
      if (ErrorOnNoConverge) assert(TrueResidual / Tolerance < 10000.0);
      return;
    }
    host_cg(Linop, src, psi);
  }

 private:
  void host_cg(LinearOperatorBase<Field> &Linop, const Field &src, Field &psi) {
    RealD cp, c, a, d, b, ssq;
    Field p(src.Grid()), mmp(src.Grid()), r(src.Grid());
    p.Checkerboard() = mmp.Checkerboard() = r.Checkerboard() = src.Checkerboard();
    ssq = norm2(src);
    RealD guess = norm2(psi);
    assert(std::isnan(guess) == 0);
    if (guess == 0.0) {
      r = src;
      p = r;
      a = ssq;
    } else {
      Linop.HermOpAndNorm(psi, mmp, d, b);
      r = src - mmp;
      p = r;
      a = norm2(p);
    }
    cp = a;
    if (ssq == 0.) {
      psi = Zero();
      IterationsToComplete = 1;
      TrueResidual = 0.;
      return;
    }
    const RealD rsq = Tolerance * Tolerance * ssq;
    if (cp <= rsq) {
      TrueResidual = std::sqrt(a / ssq);
      IterationsToComplete = 0;
      return;
    }
    for (int k = 1; k <= MaxIterations; k++) {
      c = cp;
      Linop.HermOp(p, mmp);
      d = innerProduct(p, mmp).real();
      a = c / d;
      cp = axpy_norm(r, -a, mmp, r);
      b = cp / c;
      axpy(psi, a, p, psi);
      axpby(p, b, 1.0, p, r);
      if (cp <= rsq) {
        Linop.HermOp(psi, mmp);
        p = mmp - src;
        TrueResidual = std::sqrt(norm2(p)) / std::sqrt(norm2(src));
        IterationsToComplete = k;
        std::cout << GridLogMessage << "ConjugateGradient Converged on iteration " << k << "\tComputed residual "
                  << std::sqrt(cp / ssq) << "\tTrue residual " << TrueResidual << "\tTarget " << Tolerance
                  << std::endl;
        if (ErrorOnNoConverge) assert(TrueResidual / Tolerance < 10000.0);
        return;
      }
    }
    std::cout << GridLogMessage << "ConjugateGradient did NOT converge " << MaxIterations << std::endl;
    if (ErrorOnNoConverge) assert(0);
    IterationsToComplete = MaxIterations;
  }
};

// Grid: MixedPrecisionConjugateGradient<FieldD,FieldF>(tol, maxinner, maxouter, sp_grid, Linop_f, Linop_d).
// Defect correction with FP32 inner CG on the device; Linop_f must wrap the same action in single
// precision (Grid builds it from a precisionChange'd gauge field; libgx keeps both precisions of
// the one action).
template <class FieldD, class FieldF>
class MixedPrecisionConjugateGradient {
 public:
  RealD Tolerance, InnerTolerance;
  Integer MaxInnerIterations, MaxOuterIterations;
  GridBase *SinglePrecGrid;
  LinearOperatorBase<FieldF> &Linop_f;
  LinearOperatorBase<FieldD> &Linop_d;
  Integer TotalInnerIterations = 0, TotalOuterIterations = 0, TotalFinalStepIterations = 0;
  RealD TrueResidual = 0;
  bool InnerHalf = false;  // no Grid counterpart: FP16-storage inner solves instead of FP32
  RealD OuterLoopNormMult = 100.;
  MixedPrecisionConjugateGradient(RealD tol, Integer maxinnerit, Integer maxouterit, GridBase *sp_grid,
                                  LinearOperatorBase<FieldF> &lf, LinearOperatorBase<FieldD> &ld)
      : Tolerance(tol), InnerTolerance(tol), MaxInnerIterations(maxinnerit), MaxOuterIterations(maxouterit),
        SinglePrecGrid(sp_grid), Linop_f(lf), Linop_d(ld) {}
  void operator()(const FieldD &src_d_in, FieldD &sol_d) {
    gx_op op = Linop_d.schur_handle();
    if (!detail::same_action(op, Linop_f.schur_handle())) {
      host_loop(src_d_in, sol_d);
      return;
    }
    sol_d.Checkerboard() = src_d_in.Checkerboard();
    gx_solver_stats st{};
    const int rc = gx_mixed_cg(op, src_d_in.f, sol_d.f, Tolerance, InnerTolerance, InnerHalf ? GX_F16 : GX_F32,
                               MaxInnerIterations, MaxOuterIterations, &st);
    TotalInnerIterations = st.inner_iterations;
    TotalOuterIterations = st.outer_iterations;
    TotalFinalStepIterations = st.final_iterations;
    TrueResidual = st.true_residual;
    if (rc != GX_ENOCONV) detail::check(rc, "MixedPrecisionConjugateGradient");
    std::cout << GridLogMessage << "MixedPrecisionConjugateGradient: Inner CG iterations " << TotalInnerIterations
              << " Restarts " << TotalOuterIterations << " Final CG iterations " << TotalFinalStepIterations
              << " True residual " << TrueResidual << std::endl;
  }

 private:
  // Operators without a device handle: Grid's loop (ConjugateGradientMixedPrec.h:71-165) with the
  // facade's ConjugateGradient in each precision and precisionChange between them.
  void host_loop(const FieldD &src_d_in, FieldD &sol_d) {
    TotalInnerIterations = 0;
    const int cb = src_d_in.Checkerboard();
    sol_d.Checkerboard() = cb;
    const RealD stop = norm2(src_d_in) * Tolerance * Tolerance;
    GridBase *gd = src_d_in.Grid();
    FieldD tmp_d(gd), src_d(gd);
    tmp_d.Checkerboard() = cb;
    src_d = src_d_in;
    RealD inner_tol = InnerTolerance;
    FieldF src_f(SinglePrecGrid), sol_f(SinglePrecGrid);
    src_f.Checkerboard() = cb;
    sol_f.Checkerboard() = cb;
    ConjugateGradient<FieldF> CG_f(inner_tol, MaxInnerIterations, false);
    for (TotalOuterIterations = 0; TotalOuterIterations < MaxOuterIterations; TotalOuterIterations++) {
      Linop_d.HermOp(sol_d, tmp_d);
      const RealD norm = axpy_norm(src_d, -1., tmp_d, src_d_in);  // residual
      if (norm < OuterLoopNormMult * stop) break;
      while (norm * inner_tol * inner_tol < stop) inner_tol *= 2;
      precisionChange(src_f, src_d);
      sol_f = Zero();
      sol_f.Checkerboard() = cb;
      CG_f.Tolerance = inner_tol;
      CG_f(Linop_f, src_f, sol_f);
      TotalInnerIterations += CG_f.IterationsToComplete;
      precisionChange(tmp_d, sol_f);
      axpy(sol_d, 1.0, tmp_d, sol_d);
    }
    ConjugateGradient<FieldD> CG_d(Tolerance, MaxInnerIterations);
    CG_d(Linop_d, src_d_in, sol_d);
    TotalFinalStepIterations = CG_d.IterationsToComplete;
    TrueResidual = CG_d.TrueResidual;
  }
};

// Grid: MixedPrecisionConjugateGradientBatched<FieldD,FieldF>(tol, maxinner, maxouter, maxpatchup, sp_grid,
// Linop_f, Linop_d, updateResidual) (ConjugateGradientMixedPrecBatched.h:37-213): defect correction
// for a batch of sources with one inner tolerance for the batch.  When Linop_f is the same action
// as Linop_d (a Schur operator over libgx) the batch runs on the device (gx_mixed_cg_batched): on a
// Wilson operator as one multi-RHS solve sharing every link load, otherwise source by source;
// InnerHalf = true keeps the inner vectors in FP16 storage.  Other operators, and updateResidual =
// false, run Grid's loop here with the facade's ConjugateGradient.
template <class FieldD, class FieldF>
class MixedPrecisionConjugateGradientBatched {
 public:
  RealD Tolerance, InnerTolerance;
  Integer MaxInnerIterations, MaxOuterIterations, MaxPatchupIterations;
  GridBase *SinglePrecGrid;
  RealD OuterLoopNormMult = 100.;
  LinearOperatorBase<FieldF> &Linop_f;
  LinearOperatorBase<FieldD> &Linop_d;
  bool updateResidual;
  bool InnerHalf = false;  // no Grid counterpart: FP16-storage inner solves instead of FP32
  Integer TotalOuterIterations = 0;
  std::vector<Integer> TotalInnerIterations, TotalFinalStepIterations;
  MixedPrecisionConjugateGradientBatched(RealD tol, Integer maxinnerit, Integer maxouterit, Integer maxpatchit,
                                         GridBase *sp_grid, LinearOperatorBase<FieldF> &lf,
                                         LinearOperatorBase<FieldD> &ld, bool update_residual = true)
      : Tolerance(tol), InnerTolerance(tol), MaxInnerIterations(maxinnerit), MaxOuterIterations(maxouterit),
        MaxPatchupIterations(maxpatchit), SinglePrecGrid(sp_grid), Linop_f(lf), Linop_d(ld),
        updateResidual(update_residual) {}
  void operator()(const FieldD &src_d_in, FieldD &sol_d) {
    std::vector<FieldD> srcs, sols;
    srcs.emplace_back(src_d_in.Grid());
    sols.emplace_back(sol_d.Grid());
    srcs[0] = src_d_in;
    sols[0] = sol_d;
    (*this)(srcs, sols);
    sol_d = sols[0];
  }
  void operator()(const std::vector<FieldD> &src_d_in, std::vector<FieldD> &sol_d) {
    const int n = (int)src_d_in.size();
    assert(n == (int)sol_d.size() && n >= 1);
    gx_op op = Linop_d.schur_handle();
    if (!updateResidual || !detail::same_action(op, Linop_f.schur_handle())) {
      host_loop(src_d_in, sol_d);
      return;
    }
    std::vector<gx_field> hs(n), hx(n);
    for (int i = 0; i < n; i++) {
      sol_d[i].Checkerboard() = src_d_in[i].Checkerboard();
      hs[i] = src_d_in[i].f;
      hx[i] = sol_d[i].f;
    }
    std::vector<gx_solver_stats> st(n);
    const int rc = gx_mixed_cg_batched(op, n, hs.data(), hx.data(), Tolerance, InnerTolerance,
                                       InnerHalf ? GX_F16 : GX_F32, MaxInnerIterations, MaxOuterIterations,
                                       MaxPatchupIterations, st.data());
    if (rc != GX_ENOCONV) detail::check(rc, "MixedPrecisionConjugateGradientBatched");
    TotalInnerIterations.assign(n, 0);
    TotalFinalStepIterations.assign(n, 0);
    TotalOuterIterations = st[0].outer_iterations;
    for (int i = 0; i < n; i++) {
      TotalInnerIterations[i] = st[i].inner_iterations;
      TotalFinalStepIterations[i] = st[i].final_iterations;
      std::cout << GridLogMessage << "MixedPrecisionConjugateGradientBatched: solve " << i << " Inner CG iterations "
                << TotalInnerIterations[i] << " Restarts " << st[i].outer_iterations << " Final CG iterations "
                << TotalFinalStepIterations[i] << " true residual " << st[i].true_residual << std::endl;
    }
  }

 private:
  // Grid's loop (ConjugateGradientMixedPrecBatched.h:77-202) over the facade's BLAS1 and CG
  void host_loop(const std::vector<FieldD> &src_d_in, std::vector<FieldD> &sol_d) {
    const int n = (int)src_d_in.size();
    const int cb = src_d_in[0].Checkerboard();
    GridBase *gd = src_d_in[0].Grid();
    std::vector<RealD> src_norm(n), norm(n, 0.0), stop(n);
    std::vector<FieldD> src_d;
    std::vector<FieldF> src_f, sol_f;
    FieldD tmp_d(gd);
    tmp_d.Checkerboard() = cb;
    for (int i = 0; i < n; i++) {
      sol_d[i].Checkerboard() = cb;
      src_norm[i] = norm2(src_d_in[i]);
      stop[i] = src_norm[i] * Tolerance * Tolerance;
      src_d.emplace_back(gd);
      src_d[i] = src_d_in[i];
      src_f.emplace_back(SinglePrecGrid);
      sol_f.emplace_back(SinglePrecGrid);
      src_f[i].Checkerboard() = cb;
      sol_f[i].Checkerboard() = cb;
    }
    TotalInnerIterations.assign(n, 0);
    TotalFinalStepIterations.assign(n, 0);
    RealD inner_tol = InnerTolerance;
    ConjugateGradient<FieldF> CG_f(inner_tol, MaxInnerIterations, false);
    for (TotalOuterIterations = 0; TotalOuterIterations < MaxOuterIterations; TotalOuterIterations++) {
      bool all = true;
      for (int i = 0; i < n; i++) {
        Linop_d.HermOp(sol_d[i], tmp_d);
        norm[i] = axpy_norm(src_d[i], -1., tmp_d, src_d_in[i]);
        precisionChange(src_f[i], src_d[i]);
        sol_f[i] = Zero();
        sol_f[i].Checkerboard() = cb;
        if (norm[i] > OuterLoopNormMult * stop[i]) all = false;
      }
      if (all) break;
      if (updateResidual) {
        const RealD nmax = *std::max_element(norm.begin(), norm.end());
        const RealD smax = *std::max_element(stop.begin(), stop.end());
        while (nmax * inner_tol * inner_tol < smax) inner_tol *= 2;
        CG_f.Tolerance = inner_tol;
      }
      for (int i = 0; i < n; i++) {
        CG_f(Linop_f, src_f[i], sol_f[i]);
        TotalInnerIterations[i] += CG_f.IterationsToComplete;
        precisionChange(tmp_d, sol_f[i]);
        axpy(sol_d[i], 1.0, tmp_d, sol_d[i]);
      }
    }
    for (int i = 0; i < n; i++) {
      ConjugateGradient<FieldD> CG_d(Tolerance, MaxPatchupIterations);
      CG_d(Linop_d, src_d_in[i], sol_d[i]);
      TotalFinalStepIterations[i] += CG_d.IterationsToComplete;
    }
  }
};

// Grid: ConjugateGradientReliableUpdate<FieldD,FieldF>(tol, maxit, delta, sp_grid, Linop_f, Linop_d,
// err_on_no_conv)(src, psi), ConjugateGradientReliableUpdate.h:36-269.  The sloppy Krylov vectors
// are FieldF's precision (FP32); SloppyHalf = true keeps them in FP16 storage instead.  The final
// FP64 cleanup CG always runs (Grid's DoFinalCleanup default).  setFallbackLinop (:72-75,255-259):
// with SloppyHalf the FP16 vectors move to FP32 storage once the sloppy |r|^2 drops below the
// transition tolerance; with FP32 sloppy vectors the fallback operator is the same single-precision
// operator (there is no compressed-communication form to leave), so the call changes nothing.
template <class FieldD, class FieldF>
class ConjugateGradientReliableUpdate {
 public:
  bool ErrorOnNoConverge;
  RealD Tolerance;
  Integer MaxIterations;
  Integer IterationsToComplete = 0, ReliableUpdatesPerformed = 0, IterationsToCleanup = 0;
  bool DoFinalCleanup = true;
  bool SloppyHalf = false;
  LinearOperatorBase<FieldF> &Linop_f;
  LinearOperatorBase<FieldD> &Linop_d;
  GridBase *SinglePrecGrid;
  RealD Delta;
  RealD TrueResidual = 0;
  LinearOperatorBase<FieldF> *Linop_fallback = nullptr;
  RealD fallback_transition_tol = 0;
  void setFallbackLinop(LinearOperatorBase<FieldF> &lf, const RealD tol) {
    Linop_fallback = &lf;
    fallback_transition_tol = tol;
  }
  ConjugateGradientReliableUpdate(RealD tol, Integer maxit, RealD delta, GridBase *sp_grid,
                                  LinearOperatorBase<FieldF> &lf, LinearOperatorBase<FieldD> &ld,
                                  bool err_on_no_conv = true)
      : ErrorOnNoConverge(err_on_no_conv), Tolerance(tol), MaxIterations(maxit), Linop_f(lf), Linop_d(ld),
        SinglePrecGrid(sp_grid), Delta(delta) {
    assert(Delta > 0. && Delta < 1. && "Expect  0 < Delta < 1");
  }
  void operator()(const FieldD &src, FieldD &psi) {
    gx_op op = Linop_d.schur_handle();
    if (!detail::same_action(op, Linop_f.schur_handle()) ||
        (Linop_fallback && !detail::same_action(op, Linop_fallback->schur_handle()))) {
      host_loop(src, psi);
      return;
    }
    psi.Checkerboard() = src.Checkerboard();
    gx_solver_stats st{};
    const int rc = SloppyHalf && Linop_fallback
                       ? gx_reliable_update_cg_fallback(op, src.f, psi.f, Tolerance, MaxIterations, Delta, GX_F16,
                                                        GX_F32, fallback_transition_tol, &st)
                       : gx_reliable_update_cg(op, src.f, psi.f, Tolerance, MaxIterations, Delta,
                                               SloppyHalf ? GX_F16 : GX_F32, &st);
    IterationsToComplete = st.inner_iterations;
    ReliableUpdatesPerformed = st.outer_iterations;
    IterationsToCleanup = st.final_iterations;
    TrueResidual = st.true_residual;
    if (rc == GX_ENOCONV) {
      std::cout << GridLogMessage << "ConjugateGradientReliableUpdate did NOT converge" << std::endl;
      if (ErrorOnNoConverge) assert(0);
      return;
    }
    detail::check(rc, "ConjugateGradientReliableUpdate");
    std::cout << GridLogMessage << "ConjugateGradientReliableUpdate Converged on iteration " << IterationsToComplete
              << " after " << ReliableUpdatesPerformed << " reliable updates; cleanup iterations "
              << IterationsToCleanup << "; true residual " << TrueResidual << std::endl;
  }

 private:
  // Operators without a device handle: Grid's loop (ConjugateGradientReliableUpdate.h:77-269) over the
  // facade's BLAS1 with FieldF sloppy vectors (SloppyHalf does not apply), fallback operator included.
  void host_loop(const FieldD &src, FieldD &psi) {
    const int cb = src.Checkerboard();
    psi.Checkerboard() = cb;
    GridBase *gd = src.Grid();
    LinearOperatorBase<FieldF> *Linop_f_use = &Linop_f;
    bool using_fallback = false;
    FieldD p(gd), mmp(gd), r(gd);
    p.Checkerboard() = mmp.Checkerboard() = r.Checkerboard() = cb;
    RealD cp, c, a, d, b, ssq, qq;
    assert(!std::isnan(norm2(psi)));
    Linop_d.HermOpAndNorm(psi, mmp, d, b);
    r = src - mmp;
    p = r;
    a = norm2(p);
    cp = a;
    ssq = norm2(src);
    const RealD rsq = Tolerance * Tolerance * ssq;
    IterationsToComplete = ReliableUpdatesPerformed = IterationsToCleanup = 0;
    if (cp <= rsq) {
      TrueResidual = std::sqrt(cp / ssq);
      return;
    }
    FieldF r_f(SinglePrecGrid), psi_f(SinglePrecGrid), p_f(SinglePrecGrid), mmp_f(SinglePrecGrid);
    r_f.Checkerboard() = psi_f.Checkerboard() = p_f.Checkerboard() = mmp_f.Checkerboard() = cb;
    precisionChange(r_f, r);
    psi_f = Zero();
    psi_f.Checkerboard() = cb;
    p_f = r_f;
    RealD MaxResidSinceLastRelUp = cp;
    int k, l = 0;
    for (k = 1; k <= MaxIterations; k++) {
      c = cp;
      Linop_f_use->HermOpAndNorm(p_f, mmp_f, d, qq);
      a = c / d;
      cp = axpy_norm(r_f, -a, mmp_f, r_f);
      b = cp / c;
      axpy(psi_f, a, p_f, psi_f);
      if (cp > MaxResidSinceLastRelUp) MaxResidSinceLastRelUp = cp;
      if (cp <= rsq) {
        precisionChange(mmp, psi_f);
        psi = psi + mmp;
        IterationsToComplete = k;
        ReliableUpdatesPerformed = l;
        if (DoFinalCleanup) {
          ConjugateGradient<FieldD> CG(Tolerance, MaxIterations, ErrorOnNoConverge);
          CG(Linop_d, src, psi);
          IterationsToCleanup = CG.IterationsToComplete;
          TrueResidual = CG.TrueResidual;
        } else {
          Linop_d.HermOpAndNorm(psi, mmp, d, qq);
          p = mmp - src;
          TrueResidual = std::sqrt(norm2(p) / ssq);
        }
        return;
      } else if (cp < Delta * MaxResidSinceLastRelUp) {  // reliable update
        precisionChange(mmp, psi_f);
        psi = psi + mmp;
        Linop_d.HermOpAndNorm(psi, mmp, d, qq);
        r = src - mmp;
        psi_f = Zero();
        psi_f.Checkerboard() = cb;
        precisionChange(r_f, r);
        cp = norm2(r);
        MaxResidSinceLastRelUp = cp;
        b = cp / c;
        l++;
      }
      axpy(p_f, b, p_f, r_f);  // p = b p + r
      if (!using_fallback && Linop_fallback != nullptr && cp < fallback_transition_tol) {
        Linop_f_use = Linop_fallback;
        using_fallback = true;
      }
    }
    IterationsToComplete = MaxIterations;
    ReliableUpdatesPerformed = l;
    std::cout << GridLogMessage << "ConjugateGradientReliableUpdate did NOT converge" << std::endl;
    if (ErrorOnNoConverge) assert(0);
  }
};

// Grid: MultiShiftFunction (Grid/algorithms/approx/MultiShiftFunction.h:33-62) without the Remez
// initialisation (AlgRemez needs GMP): the caller fills poles, residues, tolerances and norm.  The
// (n, lo, hi) constructor also sets order = n, which Grid leaves to the caller.
class MultiShiftFunction {
 public:
  int order = 0;
  std::vector<RealD> poles, residues, tolerances;
  RealD norm = 0, lo = 0, hi = 0;
  MultiShiftFunction() = default;
  MultiShiftFunction(int n, RealD _lo, RealD _hi)
      : order(n), poles(n), residues(n), tolerances(n), lo(_lo), hi(_hi) {}
  // norm + sum_i residues[i] / (x + poles[i])
  RealD approx(RealD x) const {
    RealD a = norm;
    for (int i = 0; i < order; i++) a += residues[i] / (x + poles[i]);
    return a;
  }
};

// Grid: ConjugateGradientMultiShift<Field>(maxit, shifts), ConjugateGradientMultiShift.h:40-343.  On
// a Schur operator over a libgx action: gx_cg_multishift for (HermOp + poles[s]) psi_s = src.  Like
// Grid it prints, but does not assert, when a shift fails to converge.
template <class Field>
class ConjugateGradientMultiShift : public OperatorFunction<Field> {
 public:
  Integer MaxIterations;
  Integer IterationsToComplete = 0;
  std::vector<int> IterationsToCompleteShift;
  int verbose = 1;
  MultiShiftFunction shifts;
  std::vector<RealD> TrueResidualShift;
  ConjugateGradientMultiShift(Integer maxit, const MultiShiftFunction &_shifts)
      : MaxIterations(maxit), shifts(_shifts) {
    IterationsToCompleteShift.resize(shifts.order);
    TrueResidualShift.resize(shifts.order);
  }
  // psi = norm src + sum_s residues[s] (HermOp + poles[s])^-1 src
  void operator()(LinearOperatorBase<Field> &Linop, const Field &src, Field &psi) override {
    std::vector<Field> results;
    results.reserve(shifts.order);
    for (int s = 0; s < shifts.order; s++) results.emplace_back(src.Grid());
    (*this)(Linop, src, results, psi);
  }
  void operator()(LinearOperatorBase<Field> &Linop, const Field &src, std::vector<Field> &results, Field &psi) {
    (*this)(Linop, src, results);
    psi.Checkerboard() = src.Checkerboard();
    psi = shifts.norm * src;
    for (int s = 0; s < shifts.order; s++) psi = psi + shifts.residues[s] * results[s];
  }
  void operator()(LinearOperatorBase<Field> &Linop, const Field &src, std::vector<Field> &psi) {
    gx_op op = Linop.schur_handle();
    const int n = shifts.order;
    assert((int)psi.size() == n && (int)shifts.poles.size() == n && (int)shifts.tolerances.size() == n);
    // the device solver takes FP64 fields and at most 32 poles; Grid takes any, so the rest run here
    if (!op || Field::Prec != GX_F64 || n > 32) {
      host_loop(Linop, src, psi);
      return;
    }
    std::vector<gx_field> h(n);
    for (int s = 0; s < n; s++) {
      psi[s].Checkerboard() = src.Checkerboard();
      h[s] = psi[s].f;
    }
    std::vector<gx_solver_stats> st(n);
    const int rc = gx_cg_multishift(op, src.f, n, shifts.poles.data(), shifts.tolerances.data(), h.data(),
                                    MaxIterations, st.data());
    if (rc != GX_ENOCONV) detail::check(rc, "ConjugateGradientMultiShift");
    IterationsToCompleteShift.assign(n, 0);
    TrueResidualShift.assign(n, 0.0);
    IterationsToComplete = 0;
    for (int s = 0; s < n; s++) {
      IterationsToCompleteShift[s] = st[s].iterations;
      TrueResidualShift[s] = st[s].true_residual;
      IterationsToComplete = std::max<Integer>(IterationsToComplete, st[s].iterations);
      if (verbose)
        std::cout << GridLogMessage << "CGMultiShift: shift[" << s << "] true residual " << TrueResidualShift[s]
                  << " after " << st[s].iterations << " iterations" << std::endl;
    }
    if (rc == GX_ENOCONV) std::cout << GridLogMessage << "CG multi shift did not converge" << std::endl;
  }

 private:
  // Any other LinearOperatorBase: Grid's loop (ConjugateGradientMultiShift.h:130-338) over the
  // facade's device BLAS1, one operator call per iteration through the user's HermOp.
  void host_loop(LinearOperatorBase<Field> &Linop, const Field &src, std::vector<Field> &psi) {
    const int n = shifts.order;
    const std::vector<RealD> &mass = shifts.poles, &mres = shifts.tolerances;
    for (int s = 0; s < n; s++) assert(mass[s] >= mass[0]);
    GridBase *grid = src.Grid();
    std::vector<Field> ps;
    for (int s = 0; s < n; s++) {
      ps.emplace_back(grid);
      ps[s].Checkerboard() = src.Checkerboard();
      psi[s].Checkerboard() = src.Checkerboard();
    }
    Field r(grid), p(grid), mmp(grid), tmp(grid);
    r.Checkerboard() = p.Checkerboard() = mmp.Checkerboard() = tmp.Checkerboard() = src.Checkerboard();
    std::vector<RealD> bs(n), rsq(n), z0(n, 1.0), z1(n, 1.0);  // z1 = z[s][iz], z0 = z[s][1-iz]
    std::vector<int> conv(n, 0);
    IterationsToCompleteShift.assign(n, 0);
    TrueResidualShift.assign(n, 0.0);
    RealD cp = norm2(src);
    if (cp == 0.0) {
      for (int s = 0; s < n; s++) {
        psi[s] = Zero();
        IterationsToCompleteShift[s] = 1;
      }
      return;
    }
    for (int s = 0; s < n; s++) {
      rsq[s] = cp * mres[s] * mres[s];
      ps[s] = src;
    }
    r = src;
    p = src;
    RealD d, qq;
    Linop.HermOpAndNorm(p, mmp, d, qq);
    axpy(mmp, mass[0], p, mmp);
    d += norm2(p) * mass[0];
    RealD b = -cp / d, c;
    bs[0] = b;
    for (int s = 1; s < n; s++) {
      z1[s] = 1.0 / (1.0 - b * (mass[s] - mass[0]));
      bs[s] = b * z1[s];
    }
    c = axpy_norm(r, b, mmp, r);
    for (int s = 0; s < n; s++) axpby(psi[s], 0.0, -bs[s], src, src);
    int k;
    for (k = 1; k <= MaxIterations; k++) {
      const RealD a = c / cp;
      axpy(p, a, p, r);
      for (int s = 0; s < n; s++) {
        if (conv[s]) continue;
        if (s == 0) axpy(ps[0], a, ps[0], r);
        else axpby(ps[s], z1[s], a * z1[s] * bs[s] / (z0[s] * b), r, ps[s]);
      }
      cp = c;
      Linop.HermOp(p, mmp);
      d = innerProduct(p, mmp).real();
      axpy(mmp, mass[0], p, mmp);
      d += norm2(p) * mass[0];
      const RealD bp = b;
      b = -cp / d;
      c = axpy_norm(r, b, mmp, r);
      bs[0] = b;
      for (int s = 1; s < n; s++) {
        if (conv[s]) continue;
        const RealD zo = z1[s], zp = z0[s];  // current and previous before the toggle
        const RealD zn = zo * zp * bp / (b * a * (zp - zo) + zp * bp * (1.0 - (mass[s] - mass[0]) * b));
        z0[s] = zo;
        z1[s] = zn;
        bs[s] = b * zn / zo;
      }
      for (int s = 0; s < n; s++)
        if (!conv[s]) axpy(psi[s], -bs[s], ps[s], psi[s]);
      bool all = true;
      for (int s = 0; s < n; s++) {
        if (conv[s]) continue;
        IterationsToCompleteShift[s] = k;
        if (c * z1[s] * z1[s] < rsq[s]) conv[s] = 1;
        else all = false;
      }
      if (all) break;
    }
    IterationsToComplete = std::min<Integer>(k, MaxIterations);
    const RealD cn = norm2(src);
    for (int s = 0; s < n; s++) {
      Linop.HermOp(psi[s], mmp);
      axpy(tmp, mass[s], psi[s], mmp);
      axpy(r, -1.0, src, tmp);
      TrueResidualShift[s] = std::sqrt(norm2(r) / cn);
      if (verbose)
        std::cout << GridLogMessage << "CGMultiShift: shift[" << s << "] true residual " << TrueResidualShift[s]
                  << " after " << IterationsToCompleteShift[s] << " iterations" << std::endl;
    }
  }
};

// Grid: BlockConjugateGradient<Field>(BlockCGtype, Orthog, tol, maxit, err_on_no_conv),
// BlockConjugateGradient.h:86-111,197-207.  The block is the fifth dimension (Orthog = 0) of a
// multi-RHS Wilson operator's Schur system: CGmultiRHS runs gx_block_cg, BlockCGrQ gx_block_cgrq;
// like Grid, the other types assert for a single field.  TrueResidual is Grid's block aggregate
// sqrt(sum_b |A X_b - B_b|^2 / sum_b |B_b|^2).
enum BlockCGtype { BlockCG, BlockCGrQ, CGmultiRHS, BlockCGVec, BlockCGrQVec };

template <class Field>
class BlockConjugateGradient : public OperatorFunction<Field> {
 public:
  int blockDim;
  int Nblock = 0;
  BlockCGtype CGtype;
  bool ErrorOnNoConverge;
  RealD Tolerance;
  Integer MaxIterations;
  Integer IterationsToComplete = 0;
  Integer PrintInterval = 100;
  RealD TrueResidual = 0;
  BlockConjugateGradient(BlockCGtype cgtype, int _Orthog, RealD tol, Integer maxit, bool err_on_no_conv = true)
      : blockDim(_Orthog), CGtype(cgtype), ErrorOnNoConverge(err_on_no_conv), Tolerance(tol), MaxIterations(maxit) {}

  void operator()(LinearOperatorBase<Field> &Linop, const Field &Src, Field &Psi) override {
    gx_op op = Linop.schur_handle();
    assert(op && "BlockConjugateGradient runs on a Schur operator over a multi-RHS WilsonFermion");
    assert(blockDim == 0 && "the block direction is the fifth dimension (Orthog = 0)");
    assert((CGtype == BlockCGrQ || CGtype == CGmultiRHS) && "single-field BlockConjugateGradient: BlockCGrQ or CGmultiRHS");
    Nblock = Src.Grid()->Ls;
    Psi.Checkerboard() = Src.Checkerboard();
    std::vector<gx_solver_stats> st(Nblock);
    const int rc = CGtype == BlockCGrQ ? gx_block_cgrq(op, Src.f, Psi.f, Tolerance, MaxIterations, st.data())
                                       : gx_block_cg(op, Src.f, Psi.f, Tolerance, MaxIterations, GX_F64, st.data());
    if (rc != GX_ENOCONV) detail::check(rc, "BlockConjugateGradient");
    std::vector<double> m(2 * (size_t)Nblock * Nblock);
    detail::check(gx_slice_inner_product_matrix(Src.f, Src.f, m.data()), "sliceInnerProductMatrix");
    double rr = 0, ss = 0;
    IterationsToComplete = 0;
    for (int b = 0; b < Nblock; b++) {
      const double nb = m[2 * ((size_t)b * Nblock + b)];
      rr += st[b].true_residual * st[b].true_residual * nb;
      ss += nb;
      IterationsToComplete = std::max<Integer>(IterationsToComplete, st[b].iterations);
    }
    TrueResidual = ss > 0 ? std::sqrt(rr / ss) : 0.0;
    if (rc == GX_ENOCONV) {
      std::cout << GridLogMessage << "BlockConjugateGradient did NOT converge " << IterationsToComplete << " / "
                << MaxIterations << std::endl;
      if (ErrorOnNoConverge) assert(0);
      return;
    }
    std::cout << GridLogMessage << (CGtype == BlockCGrQ ? "BlockCGrQ" : "MultiRHS solver") << " converged in "
              << IterationsToComplete << " iterations; true residual " << TrueResidual << std::endl;
  }

  // Grid: operator()(Linop, std::vector<Field> Src, std::vector<Field> Psi) (BlockConjugateGradient.h:
  // 206-213): BlockCGrQVec only, as in Grid.  On a Schur operator (MpcDagMpc, half fields) or an
  // MdagMLinearOperator (full fields) over a libgx action with FP64 fields the whole solve runs on the
  // device (gx_block_cgrq_vec); otherwise Grid's BlockCGrQsolveVec loop (:554-741) runs here with the
  // user's HermOp field by field and the device InnerProductMatrix / MaddMatrix.
  virtual void operator()(LinearOperatorBase<Field> &Linop, const std::vector<Field> &Src, std::vector<Field> &Psi) {
    assert(CGtype == BlockCGrQVec && "BlockConjugateGradient over std::vector<Field>: BlockCGrQVec");
    Nblock = (int)Src.size();
    assert(Nblock == (int)Psi.size() && Nblock >= 1);
    for (int b = 0; b < Nblock; b++) Psi[b].Checkerboard() = Src[b].Checkerboard();
    const bool half = Src[0].Grid()->checkerboarded;
    gx_op op = half ? Linop.schur_handle() : Linop.mdagm_handle();
    if (!op || Field::Prec != GX_F64 || Nblock > 16) {
      host_vec(Linop, Src, Psi);
      return;
    }
    std::vector<gx_field> hs(Nblock), hx(Nblock);
    for (int b = 0; b < Nblock; b++) {
      hs[b] = Src[b].f;
      hx[b] = Psi[b].f;
    }
    std::vector<gx_solver_stats> st(Nblock);
    const int rc = gx_block_cgrq_vec(op, half ? GX_HERM_SCHUR : GX_HERM_MDAGM, Nblock, hs.data(), hx.data(), Tolerance,
                                     MaxIterations, st.data());
    if (rc != GX_ENOCONV) detail::check(rc, "BlockConjugateGradient(BlockCGrQVec)");
    double rr = 0, ss = 0;
    IterationsToComplete = 0;
    for (int b = 0; b < Nblock; b++) {
      const double nb = norm2(Src[b]);
      rr += st[b].true_residual * st[b].true_residual * nb;
      ss += nb;
      IterationsToComplete = std::max<Integer>(IterationsToComplete, st[b].iterations);
    }
    TrueResidual = ss > 0 ? std::sqrt(rr / ss) : 0.0;
    report(rc == GX_OK);
  }

 private:
  void report(bool converged) {
    if (!converged) {
      std::cout << GridLogMessage << "BlockConjugateGradient(rQ) did NOT converge" << std::endl;
      if (ErrorOnNoConverge) assert(0);
      return;
    }
    std::cout << GridLogMessage << "BlockCGrQ converged in " << IterationsToComplete << " iterations; true residual "
              << TrueResidual << std::endl;
  }
  static std::vector<gx_field> handles(const std::vector<Field> &v) {
    std::vector<gx_field> h(v.size());
    for (size_t b = 0; b < v.size(); b++) h[b] = v[b].f;
    return h;
  }
  static detail::DenseC ipm(const std::vector<Field> &X, const std::vector<Field> &Y) {
    const int n = (int)X.size();
    std::vector<double> m(2 * (size_t)n * n);
    detail::check(gx_inner_product_matrix(n, handles(X).data(), handles(Y).data(), m.data()), "InnerProductMatrix");
    detail::DenseC c(n);
    for (int i = 0; i < n * n; i++) c.a[i] = ComplexD(m[2 * i], m[2 * i + 1]);
    return c;
  }
  // R = Y + scale X a (MaddMatrix); Y = nullptr: MulMatrix
  static void madd(std::vector<Field> &R, const detail::DenseC &a, const std::vector<Field> &X,
                   const std::vector<Field> *Y, double scale = 1.0) {
    const int n = (int)X.size();
    std::vector<double> m(2 * (size_t)n * n);
    for (int i = 0; i < n * n; i++) {
      m[2 * i] = a.a[i].real();
      m[2 * i + 1] = a.a[i].imag();
    }
    std::vector<gx_field> hy = Y ? handles(*Y) : std::vector<gx_field>();
    detail::check(gx_madd_matrix(n, handles(R).data(), m.data(), handles(X).data(), Y ? hy.data() : nullptr, scale),
                  "MaddMatrix");
  }
  // ThinQRfact (BlockConjugateGradient.h:168-190): C = chol(R^dag R)^dag, Q = R C^-1
  static void thin_qr(std::vector<Field> &Q, detail::DenseC &C, const std::vector<Field> &R) {
    const detail::DenseC m = ipm(R, R);
    detail::DenseC h(m.n);
    for (int i = 0; i < m.n; i++)
      for (int j = 0; j < m.n; j++) h(i, j) = 0.5 * (m(i, j) + std::conj(m(j, i)));
    C = h.cholesky().adj();
    madd(Q, C.inverse(), R, nullptr);
  }
  void host_vec(LinearOperatorBase<Field> &Linop, const std::vector<Field> &B, std::vector<Field> &X) {
    const int n = Nblock;
    GridBase *g = B[0].Grid();
    std::vector<Field> tmp, Q, D, Z;
    for (int b = 0; b < n; b++) {
      tmp.emplace_back(g), Q.emplace_back(g), D.emplace_back(g), Z.emplace_back(g);
      tmp[b].Checkerboard() = Q[b].Checkerboard() = D[b].Checkerboard() = Z[b].Checkerboard() = B[b].Checkerboard();
    }
    std::vector<RealD> ssq(n);
    for (int b = 0; b < n; b++) {
      ssq[b] = norm2(B[b]);
      assert(!std::isnan(ssq[b]) && !std::isnan(norm2(X[b])));
    }
    for (int b = 0; b < n; b++) {
      Linop.HermOp(X[b], Z[b]);
      tmp[b] = B[b] - Z[b];
    }
    detail::DenseC C, S, M;
    thin_qr(Q, C, tmp);
    for (int b = 0; b < n; b++) D[b] = Q[b];
    int k;
    bool converged = false;
    for (k = 1; k <= MaxIterations && !converged; k++) {
      for (int b = 0; b < n; b++) Linop.HermOp(D[b], Z[b]);
      M = ipm(D, Z).inverse();
      madd(X, M * C, D, &X);
      madd(tmp, M, Z, &Q, -1.0);
      thin_qr(Q, S, tmp);
      madd(D, S.adj(), D, &Q);
      C = S * C;
      const detail::DenseC rr = C.adj() * C;
      double mx = 0;
      for (int b = 0; b < n; b++) mx = std::max(mx, rr(b, b).real() / ssq[b]);
      converged = mx < Tolerance * Tolerance;
    }
    IterationsToComplete = k - 1;
    double num = 0, den = 0;
    for (int b = 0; b < n; b++) {
      Linop.HermOp(X[b], Z[b]);
      num += norm2(Z[b] - B[b]);
      den += ssq[b];
    }
    TrueResidual = std::sqrt(num / den);
    report(converged);
  }
};

// Grid: SchurRedBlackDiagMooeeSolve<Field>(OperatorFunction&)(Matrix&, in, out), SchurRedBlack.h:236-436
template <class Field>
class SchurRedBlackDiagMooeeSolve {
 public:
  OperatorFunction<Field> &_HermitianRBSolver;
  explicit SchurRedBlackDiagMooeeSolve(OperatorFunction<Field> &solver, bool = false, bool = false)
      : _HermitianRBSolver(solver) {}
  template <class Matrix>
  void operator()(Matrix &_Matrix, const Field &in, Field &out) {
    GridBase *grid = _Matrix.RedBlackGrid();
    GridBase *fgrid = _Matrix.Grid();
    Field resid(fgrid), src_o(grid), src_e(grid), sol_o(grid);
    RedBlackSource(_Matrix, in, src_e, src_o);
    sol_o = Zero();
    sol_o.Checkerboard() = Odd;
    SchurDiagMooeeOperator<Matrix, Field> HermOpEO(_Matrix);
    _HermitianRBSolver(HermOpEO, src_o, sol_o);
    assert(sol_o.Checkerboard() == Odd);
    RedBlackSolution(_Matrix, sol_o, src_e, out);
    _Matrix.M(out, resid);
    resid = resid - in;
    RealD ns = norm2(in), nr = norm2(resid);
    std::cout << GridLogMessage << "SchurRedBlackBase solver true unprec resid " << std::sqrt(nr / ns) << std::endl;
  }
  template <class Matrix>
  void RedBlackSource(Matrix &_Matrix, const Field &src, Field &src_e, Field &src_o) {
    GridBase *grid = _Matrix.RedBlackGrid();
    Field tmp(grid), Mtmp(grid);
    pickCheckerboard(Even, src_e, src);
    pickCheckerboard(Odd, src_o, src);
    _Matrix.MooeeInv(src_e, tmp);
    assert(tmp.Checkerboard() == Even);
    _Matrix.Meooe(tmp, Mtmp);
    assert(Mtmp.Checkerboard() == Odd);
    tmp = src_o - Mtmp;
    assert(tmp.Checkerboard() == Odd);
    SchurDiagMooeeOperator<Matrix, Field> HermOpEO(_Matrix);
    HermOpEO.MpcDag(tmp, src_o);
    assert(src_o.Checkerboard() == Odd);
  }
  template <class Matrix>
  void RedBlackSolution(Matrix &_Matrix, const Field &sol_o, const Field &src_e, Field &sol) {
    GridBase *grid = _Matrix.RedBlackGrid();
    Field tmp(grid), sol_e(grid), src_e_i(grid);
    _Matrix.Meooe(sol_o, tmp);
    assert(tmp.Checkerboard() == Even);
    src_e_i = src_e - tmp;
    assert(src_e_i.Checkerboard() == Even);
    _Matrix.MooeeInv(src_e_i, sol_e);
    assert(sol_e.Checkerboard() == Even);
    setCheckerboard(sol, sol_e);
    setCheckerboard(sol, sol_o);
  }
};

}  // namespace Grid

#endif
