// ILDG gauge-configuration files (Grid/parallelIO/IldgIO.h: IldgReader::readConfiguration
// :713-936, IldgWriter::writeConfiguration :596-690; record names IldgIOtypes.h:41-52; the SciDAC
// checksum BinaryIO.h:139-193).
//
// An ILDG file is a LIME file: a sequence of records, each a 144-byte big-endian header (magic
// number 0x456789ab, format version 1 (16 bits), flags (16 bits: 0x8000 message begin, 0x4000
// message end), data length (64 bits), a 128-byte NUL-padded type string) followed by the data
// padded with zeros to a multiple of 8 bytes (the published LIME format of the USQCD c-lime
// library, which Grid links).  The records a reader needs:
//   ildg-format        XML <ildgFormat>: field su3gauge, precision 32 or 64, lx ly lz lt
//   ildg-data-lfn      the logical file name
//   ildg-binary-data   the links of every site in lexicographic order (x fastest), four
//                      directions per site, each SU(3) matrix row-major as complex numbers,
//                      big-endian IEEE32 or IEEE64
//   scidac-checksum    XML <scidacChecksum>: suma, sumb (hex)
// and, when present, Grid's own grid-format (FieldMetaData: plaquette, link_trace), the
// FieldNormMetaData record (norm2 of the field) and the USQCD scidac-record-xml <usqcdInfo>
// (plaq, linktr).  As in Grid the reader requires the format, LFN, binary and checksum records;
// checks the lattice, the field type and both SciDAC checksums (CRC-32 of each site's stored
// bytes, rotated left by the global site index mod 29 for suma and mod 31 for sumb, XORed over the
// sites); checks norm2 against FieldNormMetaData (relative 1e-5); and, when Grid or USQCD metadata
// gives them, the plaquette and link trace (1e-5), recomputed on the device.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/gx.h"

namespace gx {
void set_error(const char *fmt, ...);  // gx_last_error() text (gx_api.hip)
}

namespace {

constexpr uint32_t kLimeMagic = 0x456789abu;
constexpr size_t kLimeHeader = 144;

void put_be(uint8_t *p, uint64_t v, int bytes) {
  for (int i = bytes - 1; i >= 0; i--, v >>= 8) p[i] = (uint8_t)(v & 0xff);
}

uint64_t get_be(const uint8_t *p, int bytes) {
  uint64_t v = 0;
  for (int i = 0; i < bytes; i++) v = (v << 8) | p[i];
  return v;
}

// CRC-32 of ISO 3309 / ITU-T V.42 (reflected polynomial 0xEDB88320, initial and final value
// 0xFFFFFFFF), the checksum zlib's crc32() computes, table-driven.
uint32_t crc32_bytes(const uint8_t *p, size_t n) {
  static uint32_t table[256];
  static bool init = false;
  if (!init) {
    for (uint32_t i = 0; i < 256; i++) {
      uint32_t c = i;
      for (int k = 0; k < 8; k++) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      table[i] = c;
    }
    init = true;
  }
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < n; i++) c = table[(c ^ p[i]) & 0xff] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

uint32_t rotl(uint32_t x, unsigned r) { return r == 0 ? x : (x << r) | (x >> (32 - r)); }

// SciDAC checksums of nsite sites of site_bytes stored bytes each (global lexicographic order)
void scidac_checksum(const uint8_t *data, int64_t nsite, size_t site_bytes, uint32_t *a, uint32_t *b) {
  uint32_t sa = 0, sb = 0;
  for (int64_t s = 0; s < nsite; s++) {
    const uint32_t c = crc32_bytes(data + (size_t)s * site_bytes, site_bytes);
    sa ^= rotl(c, (unsigned)(s % 29));
    sb ^= rotl(c, (unsigned)(s % 31));
  }
  *a = sa;
  *b = sb;
}

// text of the first <tag>...</tag> (or <tag attr...>...</tag>) element, false if absent
bool xml_tag(const std::string &xml, const std::string &tag, std::string *out) {
  size_t p = 0;
  for (;;) {
    p = xml.find("<" + tag, p);
    if (p == std::string::npos) return false;
    const char next = p + 1 + tag.size() < xml.size() ? xml[p + 1 + tag.size()] : '\0';
    if (next == '>' || next == ' ' || next == '\t' || next == '\n' || next == '\r') break;
    p += 1 + tag.size();
  }
  const size_t open_end = xml.find('>', p);
  const size_t close = xml.find("</" + tag + ">", open_end);
  if (open_end == std::string::npos || close == std::string::npos) return false;
  std::string v = xml.substr(open_end + 1, close - open_end - 1);
  const size_t b = v.find_first_not_of(" \t\r\n"), e = v.find_last_not_of(" \t\r\n");
  *out = b == std::string::npos ? std::string() : v.substr(b, e - b + 1);
  return true;
}

bool xml_num(const std::string &xml, const std::string &tag, double *out) {
  std::string s;
  if (!xml_tag(xml, tag, &s)) return false;
  char *end = nullptr;
  *out = std::strtod(s.c_str(), &end);
  return end && end != s.c_str();
}

struct Record {
  std::string type;
  int mb = 0, me = 0;
  uint64_t offset = 0, length = 0;  // of the data
};

int read_records(FILE *f, std::vector<Record> *out) {
  uint8_t h[kLimeHeader];
  for (;;) {
    const size_t got = std::fread(h, 1, kLimeHeader, f);
    if (got == 0) return GX_OK;
    if (got != kLimeHeader || get_be(h, 4) != kLimeMagic) {
      gx::set_error("ildg: not a LIME record (bad header or magic number)");
      return GX_EINVAL;
    }
    Record r;
    const uint64_t flags = get_be(h + 6, 2);
    r.mb = (flags & 0x8000) ? 1 : 0;
    r.me = (flags & 0x4000) ? 1 : 0;
    r.length = get_be(h + 8, 8);
    r.type.assign(reinterpret_cast<const char *>(h + 16), strnlen(reinterpret_cast<const char *>(h + 16), 128));
    r.offset = (uint64_t)std::ftell(f);
    out->push_back(r);
    const uint64_t padded = (r.length + 7) / 8 * 8;
    if (std::fseek(f, (long)padded, SEEK_CUR) != 0) {
      gx::set_error("ildg: truncated record %s", r.type.c_str());
      return GX_EINVAL;
    }
  }
}

std::string read_text(FILE *f, const Record &r) {
  std::string s(r.length, '\0');
  std::fseek(f, (long)r.offset, SEEK_SET);
  if (r.length && std::fread(&s[0], 1, r.length, f) != r.length) return std::string();
  const size_t z = s.find('\0');
  return z == std::string::npos ? s : s.substr(0, z);
}

int write_record(FILE *f, const std::string &type, const void *data, uint64_t len, int mb, int me) {
  uint8_t h[kLimeHeader] = {};
  put_be(h, kLimeMagic, 4);
  put_be(h + 4, 1, 2);
  put_be(h + 6, (mb ? 0x8000 : 0) | (me ? 0x4000 : 0), 2);
  put_be(h + 8, len, 8);
  std::memcpy(h + 16, type.c_str(), std::min<size_t>(type.size(), 127));
  static const uint8_t zeros[8] = {};
  const uint64_t pad = (len + 7) / 8 * 8 - len;
  if (std::fwrite(h, 1, kLimeHeader, f) != kLimeHeader || (len && std::fwrite(data, 1, len, f) != len) ||
      (pad && std::fwrite(zeros, 1, pad, f) != pad)) {
    gx::set_error("ildg: write failed (record %s)", type.c_str());
    return GX_EINVAL;
  }
  return GX_OK;
}

int write_text(FILE *f, const std::string &type, const std::string &xml, int mb, int me) {
  return write_record(f, type, xml.data(), xml.size(), mb, me);
}

std::string hex32(uint32_t v) {
  char b[16];
  std::snprintf(b, sizeof b, "%x", v);
  return b;
}

}  // namespace

extern "C" {

int gx_ildg_read(gx_ctx ctx, const char *path, const int dims[4], double *U, gx_ildg_info *info) {
  if (!ctx || !path || !dims || !U) {
    gx::set_error("ildg_read: null argument");
    return GX_EINVAL;
  }
  gx_ildg_info inf{};
  FILE *f = std::fopen(path, "rb");
  if (!f) {
    gx::set_error("ildg_read: cannot open %s", path);
    return GX_EINVAL;
  }
  std::vector<Record> recs;
  int rc = read_records(f, &recs);
  std::string fmt, lfn, csum, grid, norm, usqcd;
  const Record *bin = nullptr;
  bool have_fmt = false, have_lfn = false, have_csum = false;
  for (const Record &r : recs) {
    if (r.type == "ildg-binary-data") {
      bin = &r;
    } else if (r.type == "ildg-format") {
      fmt = read_text(f, r), have_fmt = true;
    } else if (r.type == "ildg-data-lfn") {
      lfn = read_text(f, r), have_lfn = true;
    } else if (r.type == "scidac-checksum") {
      csum = read_text(f, r), have_csum = true;
    } else if (r.type == "grid-format") {
      grid = read_text(f, r);
    } else if (r.type == "FieldNormMetaData") {
      norm = read_text(f, r);
    } else if (r.type == "scidac-record-xml") {
      const std::string t = read_text(f, r);
      if (t.find("usqcdInfo") != std::string::npos) usqcd = t;
    }
  }
  auto fail = [&](const char *what) {
    gx::set_error("ildg_read: %s: %s", path, what);
    rc = GX_EINVAL;
  };
  if (rc == GX_OK && (!have_fmt || !have_lfn || !bin || !have_csum))
    fail("needs the ildg-format, ildg-data-lfn, ildg-binary-data and scidac-checksum records");
  double prec = 0, L[4] = {0, 0, 0, 0};
  std::string field;
  if (rc == GX_OK) {
    const char *names[4] = {"lx", "ly", "lz", "lt"};
    bool ok = xml_num(fmt, "precision", &prec) && xml_tag(fmt, "field", &field);
    for (int d = 0; d < 4 && ok; d++) ok = xml_num(fmt, names[d], &L[d]);
    if (!ok) fail("ildg-format lacks precision, field or lx..lt");
  }
  if (rc == GX_OK) {
    for (int d = 0; d < 4; d++) inf.dims[d] = (int)L[d];
    inf.precision = (int)prec;
    std::snprintf(inf.lfn, sizeof inf.lfn, "%s", lfn.c_str());
    for (int d = 0; d < 4 && rc == GX_OK; d++)
      if ((int)L[d] != dims[d]) fail("lattice dimensions differ from the requested ones");
    if (rc == GX_OK && field != "su3gauge") fail("field is not su3gauge");
    if (rc == GX_OK && prec != 32 && prec != 64) fail("precision is neither 32 nor 64");
  }
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const size_t word = prec == 32 ? 4 : 8, site_bytes = 4 * 9 * 2 * word;
  std::vector<uint8_t> raw;
  if (rc == GX_OK) {
    if (bin->length != (uint64_t)V * site_bytes) fail("binary record size does not match the lattice and precision");
  }
  if (rc == GX_OK) {
    raw.resize((size_t)bin->length);
    std::fseek(f, (long)bin->offset, SEEK_SET);
    if (std::fread(raw.data(), 1, raw.size(), f) != raw.size()) fail("truncated binary record");
  }
  std::fclose(f);
  if (rc != GX_OK) {
    if (info) *info = inf;
    return rc;
  }
  scidac_checksum(raw.data(), V, site_bytes, &inf.scidac_checksuma, &inf.scidac_checksumb);
  std::string sa, sb;
  if (!xml_tag(csum, "suma", &sa) || !xml_tag(csum, "sumb", &sb)) fail("scidac-checksum lacks suma or sumb");
  if (rc == GX_OK) {
    inf.scidac_checksuma_file = (uint32_t)std::strtoull(sa.c_str(), nullptr, 16);
    inf.scidac_checksumb_file = (uint32_t)std::strtoull(sb.c_str(), nullptr, 16);
  }
  // big-endian words -> host doubles, layout [site][mu][a][b] (re, im)
  const size_t nw = (size_t)V * 4 * 9 * 2;
  for (size_t i = 0; i < nw; i++) {
    const uint64_t w = get_be(raw.data() + i * word, (int)word);
    if (word == 8) {
      double d;
      std::memcpy(&d, &w, 8);
      U[i] = d;
    } else {
      const uint32_t w32 = (uint32_t)w;
      float x;
      std::memcpy(&x, &w32, 4);
      U[i] = x;
    }
  }
  double n2 = 0;
  for (size_t i = 0; i < nw; i++) n2 += U[i] * U[i];
  inf.norm2 = n2;
  if (rc == GX_OK && (inf.scidac_checksuma != inf.scidac_checksuma_file ||
                      inf.scidac_checksumb != inf.scidac_checksumb_file)) {
    gx::set_error("ildg_read: %s: SciDAC checksums %x %x computed, %x %x in the file", path, inf.scidac_checksuma,
                  inf.scidac_checksumb, inf.scidac_checksuma_file, inf.scidac_checksumb_file);
    rc = GX_EINVAL;
  }
  if (rc == GX_OK && !norm.empty()) {
    double want = 0;
    inf.has_norm = xml_num(norm, "norm2", &want) ? 1 : 0;
    inf.norm2_file = want;
    if (inf.has_norm && 0.5 * std::fabs(want - n2) / (want + n2) >= 1e-5) fail("norm2 differs from FieldNormMetaData");
  }
  // plaquette and link trace: Grid's metadata first, then USQCD's
  double plaq = 0, tr = 0;
  if (!grid.empty() && xml_num(grid, "plaquette", &plaq) && xml_num(grid, "link_trace", &tr)) inf.has_plaquette = 1;
  else if (!usqcd.empty() && xml_num(usqcd, "plaq", &plaq) && xml_num(usqcd, "linktr", &tr)) inf.has_plaquette = 1;
  if (rc == GX_OK) {
    double p = 0, t = 0;
    rc = gx_gauge_statistics(ctx, dims, U, &p, &t);
    inf.plaquette = p;
    inf.link_trace = t;
    inf.plaquette_file = plaq;
    inf.link_trace_file = tr;
    if (rc == GX_OK && inf.has_plaquette && (std::fabs(p - plaq) >= 1e-5 || std::fabs(t - tr) >= 1e-5)) {
      gx::set_error("ildg_read: %s: plaquette %.10f / link trace %.10f differ from the file's %.10f / %.10f", path, p,
                    t, plaq, tr);
      rc = GX_EINVAL;
    }
  }
  if (info) *info = inf;
  return rc;
}

int gx_ildg_write(gx_ctx ctx, const char *path, const int dims[4], const double *U, int precision, const char *lfn,
                  const char *description) {
  if (!ctx || !path || !dims || !U || (precision != 32 && precision != 64)) {
    gx::set_error("ildg_write: needs a path, the field and precision 32 or 64");
    return GX_EINVAL;
  }
  double plaq = 0, tr = 0;
  int rc = gx_gauge_statistics(ctx, dims, U, &plaq, &tr);
  if (rc != GX_OK) return rc;
  const int64_t V = (int64_t)dims[0] * dims[1] * dims[2] * dims[3];
  const size_t word = precision == 32 ? 4 : 8, site_bytes = 4 * 9 * 2 * word, nw = (size_t)V * 4 * 9 * 2;
  std::vector<uint8_t> raw(nw * word);
  for (size_t i = 0; i < nw; i++) {
    if (word == 8) {
      uint64_t w;
      std::memcpy(&w, &U[i], 8);
      put_be(raw.data() + i * 8, w, 8);
    } else {
      const float x = (float)U[i];
      uint32_t w;
      std::memcpy(&w, &x, 4);
      put_be(raw.data() + i * 4, w, 4);
    }
  }
  uint32_t sa, sb;
  scidac_checksum(raw.data(), V, site_bytes, &sa, &sb);
  std::ostringstream dims_s, fmt, info, record, file, sum;
  dims_s << dims[0] << " " << dims[1] << " " << dims[2] << " " << dims[3] << " ";
  file << "<?xml version=\"1.0\"?>\n<scidacFile>\n  <version>1.1</version>\n  <spacetime>4</spacetime>\n  <dims>"
       << dims_s.str() << "</dims>\n  <volfmt>0</volfmt>\n</scidacFile>\n";
  info.precision(17);
  info << "<?xml version=\"1.0\"?>\n<usqcdInfo>\n  <version>1.0</version>\n  <plaq>" << plaq << "</plaq>\n  <linktr>"
       << tr << "</linktr>\n  <info>" << (description ? description : "") << "</info>\n</usqcdInfo>\n";
  char date[64];
  const std::time_t now = std::time(nullptr);
  std::strftime(date, sizeof date, "%a %b %d %H:%M:%S %Y UTC", std::gmtime(&now));
  record << "<?xml version=\"1.0\"?>\n<scidacRecord>\n  <version>1.1</version>\n  <date>" << date
         << "</date>\n  <recordtype>0</recordtype>\n  <datatype>QDP_" << (precision == 32 ? "F" : "D")
         << "3_ColorMatrix</datatype>\n  <precision>" << (precision == 32 ? "F" : "D")
         << "</precision>\n  <colors>3</colors>\n  <spins>0</spins>\n  <typesize>" << 9 * 2 * word
         << "</typesize>\n  <datacount>4</datacount>\n</scidacRecord>\n";
  fmt << "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqcd.org/ildg\" "
         "xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\" "
         "xsi:schemaLocation=\"http://www.lqcd.org/ildg http://www.lqcd.org/ildg/filefmt.xsd\">\n  <version>1.0</version>\n"
         "  <field>su3gauge</field>\n  <precision>"
      << precision << "</precision>\n  <lx>" << dims[0] << "</lx>\n  <ly>" << dims[1] << "</ly>\n  <lz>" << dims[2]
      << "</lz>\n  <lt>" << dims[3] << "</lt>\n</ildgFormat>\n";
  sum << "<?xml version=\"1.0\"?>\n<scidacChecksum>\n  <version>1.0</version>\n  <suma>" << hex32(sa)
      << "</suma>\n  <sumb>" << hex32(sb) << "</sumb>\n</scidacChecksum>\n";
  FILE *f = std::fopen(path, "wb");
  if (!f) {
    gx::set_error("ildg_write: cannot create %s", path);
    return GX_EINVAL;
  }
  const std::string name = lfn ? lfn : "";
  // the record sequence of IldgWriter::writeConfiguration (IldgIO.h:675-684) without Grid's private
  // grid-format and FieldNormMetaData records: two messages, file metadata then the field record
  rc = write_text(f, "scidac-private-file-xml", file.str(), 1, 0);
  if (rc == GX_OK) rc = write_text(f, "scidac-file-xml", info.str(), 0, 1);
  if (rc == GX_OK) rc = write_text(f, "scidac-private-record-xml", record.str(), 1, 0);
  if (rc == GX_OK) rc = write_text(f, "scidac-record-xml", info.str(), 0, 0);
  if (rc == GX_OK) rc = write_text(f, "ildg-format", fmt.str(), 0, 0);
  if (rc == GX_OK) rc = write_text(f, "ildg-data-lfn", name, 0, 0);
  if (rc == GX_OK) rc = write_record(f, "ildg-binary-data", raw.data(), raw.size(), 0, 0);
  if (rc == GX_OK) rc = write_text(f, "scidac-checksum", sum.str(), 0, 1);
  if (std::fclose(f) != 0 && rc == GX_OK) {
    gx::set_error("ildg_write: closing %s failed", path);
    rc = GX_EINVAL;
  }
  return rc;
}

}  // extern "C"
