// NERSC gauge-configuration files (Grid/parallelIO/NerscIO.h:63-290; header fields
// MetaData.h:140-166; 3x2 reconstruction MetaData.h:204-218).
//
// A file is an ASCII header, "BEGIN_HEADER" ... "KEY = VALUE" lines ... "END_HEADER", followed by
// the links of every site in lexicographic order (x fastest), four directions per site, each an
// SU(3) matrix row-major as complex numbers: all three rows (DATATYPE 4D_SU3_GAUGE_3x3) or the
// first two (4D_SU3_GAUGE, third row = conj(row0 x row1)), in FLOATING_POINT IEEE32BIG, IEEE32
// (little-endian), IEEE64BIG or IEEE64 / IEEE64LITTLE.  CHECKSUM (hex) is the 32-bit sum of the
// stored words after conversion to host byte order; PLAQUETTE and LINK_TRACE are the averages
// WilsonLoops::avgPlaquette and linkTrace of the field, which a reader recomputes (on the device)
// and compares (to 1e-5 and 1e-6, as Grid).  The writer follows Grid: IEEE64BIG always, three or
// two rows, periodic boundaries.
#include <pwd.h>
#include <sys/utsname.h>
#include <unistd.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gx.h"

namespace gx {
void set_error(const char *fmt, ...);  // gx_last_error() text (gx_api.hip)
}

namespace {

std::string strip(const std::string &s) {  // removeWhitespace: drop every space and tab
  std::string o;
  for (char ch : s)
    if (ch != ' ' && ch != '\t' && ch != '\r' && ch != '\n') o += ch;
  return o;
}

bool host_little() {
  const uint16_t one = 1;
  return *reinterpret_cast<const uint8_t *>(&one) == 1;
}

void swap_words(uint8_t *p, size_t bytes, int w) {
  for (size_t i = 0; i + w <= bytes; i += w)
    for (int a = 0, b = w - 1; a < b; a++, b--) std::swap(p[i + a], p[i + b]);
}

uint32_t csum32(const uint8_t *p, size_t bytes) {
  uint32_t s = 0;
  for (size_t i = 0; i + 4 <= bytes; i += 4) {
    uint32_t w;
    std::memcpy(&w, p + i, 4);
    s += w;
  }
  return s;
}

// third row = conj(row0 x row1) (MetaData.h:212-215)
void reconstruct3(std::complex<double> m[3][3]) {
  m[2][0] = std::conj(m[0][1] * m[1][2] - m[0][2] * m[1][1]);
  m[2][1] = std::conj(m[0][2] * m[1][0] - m[0][0] * m[1][2]);
  m[2][2] = std::conj(m[0][0] * m[1][1] - m[0][1] * m[1][0]);
}

template <class F>
int parse_body(const std::vector<uint8_t> &raw, int64_t V, int rows, double *U) {
  const F *f = reinterpret_cast<const F *>(raw.data());
  for (int64_t l = 0; l < V * 4; l++) {
    std::complex<double> m[3][3] = {};
    for (int a = 0; a < rows; a++)
      for (int b = 0; b < 3; b++) {
        const int64_t k = ((l * rows + a) * 3 + b) * 2;
        m[a][b] = {(double)f[k], (double)f[k + 1]};
      }
    if (rows == 2) reconstruct3(m);
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        U[(l * 9 + a * 3 + b) * 2] = m[a][b].real();
        U[(l * 9 + a * 3 + b) * 2 + 1] = m[a][b].imag();
      }
  }
  return GX_OK;
}

void copy_str(char *dst, size_t n, const std::string &s) {
  std::snprintf(dst, n, "%s", s.c_str());
}

}  // namespace

extern "C" {

int gx_nersc_read(gx_ctx ctx, const char *path, const int dims[4], double *U, gx_nersc_info *info) {
  using gx::set_error;
  if (!ctx || !path || !dims || !U) {
    set_error("nersc_read: null argument");
    return GX_EINVAL;
  }
  std::ifstream fin(path, std::ios::binary);
  if (!fin) {
    set_error("nersc_read: cannot open %s", path);
    return GX_EINVAL;
  }
  std::string line;
  std::getline(fin, line);
  if (strip(line) != "BEGIN_HEADER") {
    set_error("nersc_read: %s does not start with BEGIN_HEADER", path);
    return GX_EINVAL;
  }
  std::map<std::string, std::string> h;
  for (;;) {
    if (!std::getline(fin, line)) {
      set_error("nersc_read: %s has no END_HEADER", path);
      return GX_EINVAL;
    }
    if (line.find("END_HEADER") != std::string::npos) break;
    const size_t eq = line.find('=');
    if (eq != std::string::npos && eq > 0) h[strip(line.substr(0, eq))] = strip(line.substr(eq + 1));
  }
  const std::streamoff data_start = fin.tellg();
  auto need = [&](const char *k) -> const std::string * {
    auto it = h.find(k);
    if (it == h.end()) {
      set_error("nersc_read: header lacks %s", k);
      return nullptr;
    }
    return &it->second;
  };
  int fdims[4];
  for (int d = 0; d < 4; d++) {
    const std::string key = "DIMENSION_" + std::to_string(d + 1);
    const std::string *v = need(key.c_str());
    if (!v) return GX_EINVAL;
    fdims[d] = std::atoi(v->c_str());
    if (fdims[d] != dims[d]) {
      set_error("nersc_read: file lattice %dx%dx%dx%d, expected %dx%dx%dx%d", fdims[0], d > 0 ? fdims[1] : 0,
                d > 1 ? fdims[2] : 0, d > 2 ? fdims[3] : 0, dims[0], dims[1], dims[2], dims[3]);
      return GX_EINVAL;
    }
  }
  const std::string *dt = need("DATATYPE"), *fp = need("FLOATING_POINT"), *cs = need("CHECKSUM"),
                    *pl = need("PLAQUETTE"), *lt = need("LINK_TRACE");
  if (!dt || !fp || !cs || !pl || !lt) return GX_EINVAL;
  int rows;
  if (*dt == "4D_SU3_GAUGE") rows = 2;
  else if (*dt == "4D_SU3_GAUGE_3x3") rows = 3;
  else {
    set_error("nersc_read: DATATYPE %s unsupported", dt->c_str());
    return GX_EINVAL;
  }
  int width;
  bool big;
  if (*fp == "IEEE32BIG") width = 4, big = true;
  else if (*fp == "IEEE32") width = 4, big = false;
  else if (*fp == "IEEE64BIG") width = 8, big = true;
  else if (*fp == "IEEE64" || *fp == "IEEE64LITTLE") width = 8, big = false;
  else {
    set_error("nersc_read: FLOATING_POINT %s unsupported", fp->c_str());
    return GX_EINVAL;
  }
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const size_t bytes = (size_t)V * 4 * rows * 3 * 2 * width;
  std::vector<uint8_t> raw(bytes);
  fin.seekg(data_start);
  fin.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)bytes);
  if ((size_t)fin.gcount() != bytes) {
    set_error("nersc_read: %s holds %lld of %zu data bytes", path, (long long)fin.gcount(), bytes);
    return GX_EINVAL;
  }
  if (big == host_little()) swap_words(raw.data(), bytes, width);  // to host byte order
  const uint32_t sum = csum32(raw.data(), bytes);
  const uint32_t hsum = (uint32_t)std::stoul(*cs, nullptr, 16);
  if (width == 4) parse_body<float>(raw, V, rows, U);
  else parse_body<double>(raw, V, rows, U);
  double plaq = 0, trace = 0;
  const int rc = gx_gauge_statistics(ctx, dims, U, &plaq, &trace);
  if (rc != GX_OK) return rc;
  const double hplaq = std::atof(pl->c_str()), htrace = std::atof(lt->c_str());
  if (info) {
    *info = gx_nersc_info{};
    for (int d = 0; d < 4; d++) info->dims[d] = fdims[d];
    info->checksum_header = hsum;
    info->checksum_data = sum;
    info->plaquette_header = hplaq;
    info->plaquette = plaq;
    info->link_trace_header = htrace;
    info->link_trace = trace;
    info->two_row = rows == 2;
    copy_str(info->floating_point, sizeof(info->floating_point), *fp);
    auto it = h.find("ENSEMBLE_ID");
    if (it != h.end()) copy_str(info->ensemble_id, sizeof(info->ensemble_id), it->second);
    it = h.find("ENSEMBLE_LABEL");
    if (it != h.end()) copy_str(info->ensemble_label, sizeof(info->ensemble_label), it->second);
    it = h.find("SEQUENCE_NUMBER");
    if (it != h.end()) info->sequence_number = std::atoi(it->second.c_str());
  }
  if (sum != hsum) {
    set_error("nersc_read: checksum mismatch in %s: data %x, header %x", path, sum, hsum);
    return GX_EINVAL;
  }
  if (!(std::fabs(plaq - hplaq) < 1e-5)) {
    set_error("nersc_read: plaquette mismatch in %s: computed %.10g, header %.10g", path, plaq, hplaq);
    return GX_EINVAL;
  }
  if (!(std::fabs(trace - htrace) < 1e-6)) {
    set_error("nersc_read: link trace mismatch in %s: computed %.10g, header %.10g", path, trace, htrace);
    return GX_EINVAL;
  }
  return GX_OK;
}

int gx_nersc_write(gx_ctx ctx, const char *path, const int dims[4], const double *U, int two_row,
                   const char *ensemble_label, const char *ensemble_id, int sequence_number) {
  using gx::set_error;
  if (!ctx || !path || !dims || !U) {
    set_error("nersc_write: null argument");
    return GX_EINVAL;
  }
  double plaq = 0, trace = 0;
  int rc = gx_gauge_statistics(ctx, dims, U, &plaq, &trace);
  if (rc != GX_OK) return rc;
  const int rows = two_row ? 2 : 3;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const size_t n = (size_t)V * 4 * rows * 3 * 2;
  std::vector<double> body(n);
  for (int64_t l = 0; l < V * 4; l++)
    for (int a = 0; a < rows; a++)
      for (int k = 0; k < 6; k++) body[(l * rows + a) * 6 + k] = U[l * 18 + a * 6 + k];
  uint8_t *raw = reinterpret_cast<uint8_t *>(body.data());
  const size_t bytes = n * sizeof(double);
  const uint32_t sum = csum32(raw, bytes);  // over host-order words, as the reader computes it
  if (host_little()) swap_words(raw, bytes, 8);  // IEEE64BIG on disk
  // creator, hardware and date (MetaData.h:120-140)
  std::string creator, hardware, date;
  if (struct passwd *pw = getpwuid(getuid())) creator = pw->pw_name;
  struct utsname un;
  if (uname(&un) == 0)
    hardware = std::string(un.nodename) + "-" + un.machine + "-" + un.sysname + "-" + un.release;
  {
    std::time_t t = std::time(nullptr);
    std::tm tm_ = *std::localtime(&t);
    std::ostringstream os;
    os << std::put_time(&tm_, "%c %Z");
    date = os.str();
  }
  std::ostringstream s;
  s << "BEGIN_HEADER\n";
  s << "HDR_VERSION = 1.0\n";
  s << "DATATYPE = " << (two_row ? "4D_SU3_GAUGE" : "4D_SU3_GAUGE_3x3") << "\n";
  s << "STORAGE_FORMAT = \n";
  for (int d = 0; d < 4; d++) s << "DIMENSION_" << d + 1 << " = " << dims[d] << "\n";
  s << "LINK_TRACE = " << std::setprecision(10) << trace << "\n";
  s << "PLAQUETTE  = " << std::setprecision(10) << plaq << "\n";
  for (int d = 0; d < 4; d++) s << "BOUNDARY_" << d + 1 << " = PERIODIC\n";
  s << "CHECKSUM = " << std::hex << std::setw(10) << sum << std::dec << "\n";
  s << "SCIDAC_CHECKSUMA = " << std::hex << std::setw(10) << 0 << std::dec << "\n";
  s << "SCIDAC_CHECKSUMB = " << std::hex << std::setw(10) << 0 << std::dec << "\n";
  s << "ENSEMBLE_ID = " << (ensemble_id ? ensemble_id : "UKQCD") << "\n";
  s << "ENSEMBLE_LABEL = " << (ensemble_label ? ensemble_label : "DWF") << "\n";
  s << "SEQUENCE_NUMBER = " << sequence_number << "\n";
  s << "CREATOR = " << creator << "\n";
  s << "CREATOR_HARDWARE = " << hardware << "\n";
  s << "CREATION_DATE = " << date << "\n";
  s << "ARCHIVE_DATE = " << date << "\n";
  s << "FLOATING_POINT = IEEE64BIG\n";
  s << "END_HEADER\n";
  std::ofstream fout(path, std::ios::binary | std::ios::trunc);
  if (!fout) {
    set_error("nersc_write: cannot open %s", path);
    return GX_EINVAL;
  }
  const std::string hdr = s.str();
  fout.write(hdr.data(), (std::streamsize)hdr.size());
  fout.write(reinterpret_cast<const char *>(raw), (std::streamsize)bytes);
  if (!fout) {
    set_error("nersc_write: write to %s failed", path);
    return GX_EINVAL;
  }
  return GX_OK;
}

}  // extern "C"
