// LIME / SciDAC / ILDG file I/O of the path's inputs and outputs (host code).
//
// LIME (the lattice QCD message encapsulation of c-lime, which the reference
// links but does not ship): a file is a sequence of records, each a 144-byte
// big-endian header -- magic 0x456789ab (4 bytes), version 1 (2 bytes), flags
// (2 bytes, 0x8000 message begin, 0x4000 message end), payload length (8
// bytes), type string (128 bytes, NUL padded) -- then the payload, padded with
// zeros to a multiple of 8 bytes.  SciDAC checksums (reference dml.c:41-53):
// per site the crc32 (zlib polynomial, initial 0) of the site's bytes as
// stored, rotated left by rank mod 29 into suma and by rank mod 31 into sumb,
// XOR-accumulated; rank is the site's lexicographic rank in the file order.
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/cvchip.h"
#include "entry_guard.h"

namespace cvchip {
void set_error(const char *fmt, ...);
}
using cvchip::set_error;

namespace {

// ------------------------------------------------------------ crc32 / DML

uint32_t crc_table[256];
bool crc_ready = false;

void crc_init() {
  for (uint32_t n = 0; n < 256; n++) {
    uint32_t c = n;
    for (int k = 0; k < 8; k++) c = (c & 1) ? 0xedb88320u ^ (c >> 1) : c >> 1;
    crc_table[n] = c;
  }
  crc_ready = true;
}

uint32_t crc32(const unsigned char *b, size_t n) {
  if (!crc_ready) crc_init();
  uint32_t c = 0xffffffffu;
  for (size_t i = 0; i < n; i++) c = crc_table[(c ^ b[i]) & 0xff] ^ (c >> 8);
  return c ^ 0xffffffffu;
}

uint32_t rotl(uint32_t w, uint32_t r) { return r ? (w << r) | (w >> (32 - r)) : w; }

struct Checksum {
  uint32_t a = 0, b = 0;
  void accum(uint64_t rank, const unsigned char *buf, size_t n) {
    const uint32_t w = crc32(buf, n);
    a ^= rotl(w, (uint32_t)(rank % 29));
    b ^= rotl(w, (uint32_t)(rank % 31));
  }
};

// ------------------------------------------------------------ byte order

void put_be64(unsigned char *p, uint64_t v) {
  for (int k = 7; k >= 0; k--, v >>= 8) p[k] = (unsigned char)(v & 0xff);
}
uint64_t get_be(const unsigned char *p, int n) {
  uint64_t v = 0;
  for (int k = 0; k < n; k++) v = (v << 8) | p[k];
  return v;
}
void put_double_be(unsigned char *p, double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  put_be64(p, u);
}
double get_double_be(const unsigned char *p) {
  const uint64_t u = get_be(p, 8);
  double d;
  memcpy(&d, &u, 8);
  return d;
}
void put_float_be(unsigned char *p, float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  for (int k = 3; k >= 0; k--, u >>= 8) p[k] = (unsigned char)(u & 0xff);
}
float get_float_be(const unsigned char *p) {
  const uint32_t u = (uint32_t)get_be(p, 4);
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// ------------------------------------------------------------ LIME records

constexpr uint32_t kMagic = 0x456789abu;

bool write_record(FILE *f, const char *type, bool mb, bool me, const void *data, uint64_t n) {
  unsigned char h[144];
  memset(h, 0, sizeof(h));
  h[0] = (kMagic >> 24) & 0xff;
  h[1] = (kMagic >> 16) & 0xff;
  h[2] = (kMagic >> 8) & 0xff;
  h[3] = kMagic & 0xff;
  h[4] = 0;
  h[5] = 1;                                     // version 1
  h[6] = (unsigned char)((mb ? 0x80 : 0) | (me ? 0x40 : 0));
  put_be64(h + 8, n);
  strncpy((char *)h + 16, type, 127);
  if (fwrite(h, 1, 144, f) != 144) return false;
  if (n && fwrite(data, 1, n, f) != n) return false;
  static const unsigned char zero[8] = {0};
  const size_t pad = (size_t)((8 - n % 8) % 8);
  return pad == 0 || fwrite(zero, 1, pad, f) == pad;
}

struct Record {
  std::string type;
  uint64_t offset, bytes;
  bool mb, me;
};

// the records of a LIME file (headers only); false if it is not one
bool scan(FILE *f, std::vector<Record> &recs) {
  recs.clear();
  if (fseek(f, 0, SEEK_END) != 0) return false;
  const long long size = ftell(f);
  long long pos = 0;
  while (pos + 144 <= size) {
    unsigned char h[144];
    if (fseek(f, pos, SEEK_SET) != 0 || fread(h, 1, 144, f) != 144) return false;
    if (get_be(h, 4) != kMagic) return false;
    Record r;
    r.bytes = get_be(h + 8, 8);
    r.mb = (h[6] & 0x80) != 0;
    r.me = (h[6] & 0x40) != 0;
    char t[129];
    memcpy(t, h + 16, 128);
    t[128] = 0;
    r.type = t;
    r.offset = (uint64_t)pos + 144;
    if (r.bytes > (uint64_t)size || r.offset + r.bytes > (uint64_t)size) return false;   // no wrap-around
    recs.push_back(r);
    pos = (long long)(r.offset + r.bytes + (8 - r.bytes % 8) % 8);
  }
  return pos == size && !recs.empty();
}

bool read_payload(FILE *f, const Record &r, std::vector<unsigned char> &buf) {
  buf.resize(r.bytes);
  return fseek(f, (long)r.offset, SEEK_SET) == 0 && fread(buf.data(), 1, r.bytes, f) == r.bytes;
}

// suma/sumb of a scidac-checksum record's XML (0 if absent)
bool parse_checksum(const std::vector<unsigned char> &xml, uint32_t &a, uint32_t &b) {
  const std::string s(xml.begin(), xml.end());
  const size_t pa = s.find("<suma>"), pb = s.find("<sumb>");
  if (pa == std::string::npos || pb == std::string::npos) return false;
  a = (uint32_t)strtoul(s.c_str() + pa + 6, nullptr, 16);
  b = (uint32_t)strtoul(s.c_str() + pb + 6, nullptr, 16);
  return true;
}

std::string checksum_xml(const Checksum &c) {
  char m[512];
  snprintf(m, sizeof(m),
           "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<scidacChecksum>\n  <version>1.0</version>\n"
           "  <suma>%08x</suma>\n  <sumb>%08x</sumb>\n</scidacChecksum>",
           c.a, c.b);
  return m;
}

bool check_dims(const int dims[4]) {
  for (int m = 0; m < 4; m++)
    if (dims[m] < 1) return false;
  return true;
}

}  // namespace

// ------------------------------------------------------------ contractions

// write_lime_contraction (reference contractions_io.c:355-420, the LEMON
// variant; the serial LIME variant's data writer is not implemented there):
// record "cvc_contraction-format" (XML), "scidac-binary-data" (per site in
// t, x, y, z order, z fastest: N complex numbers, big-endian, 64 or 32 bit),
// "scidac-checksum".
extern "C" int cvchip_lime_write_contraction(const char *filename, const double *s, int N, int prec,
                                             const int dims[4], const char *type, int nconf, int source,
                                             uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_write_contraction", [&]() -> int {
    if (!filename || !s || !dims || !type) {
      set_error("NULL argument to cvchip_lime_write_contraction");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || N < 1 || (prec != 64 && prec != 32)) {
      set_error("invalid contraction format (N %d, precision %d)", N, prec);
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "wb");
    if (!f) {
      set_error("could not open %s for writing", filename);
      return CVCHIP_ERR_ARG;
    }
    const time_t now = time(nullptr);
    char msg[2048];
    snprintf(msg, sizeof(msg),
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<cvcFormat>\n<type>%s</type>\n<precision>%d</precision>\n"
             "<components>%d</components>\n<lx>%d</lx>\n<ly>%d</ly>\n<lz>%d</lz>\n<lt>%d</lt>\n<nconf>%d</nconf>\n"
             "<source>%d</source>\n</cvcFormat>\ndate %s",
             type, prec, N, dims[1], dims[2], dims[3], dims[0], nconf, source, ctime(&now));
    bool ok = write_record(f, "cvc_contraction-format", true, true, msg, strlen(msg));
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    const size_t w = prec == 64 ? 8 : 4, site_bytes = (size_t)2 * N * w;
    std::vector<unsigned char> data(V * site_bytes);
    Checksum ck;
    for (uint64_t ix = 0; ix < V; ix++) {
      unsigned char *p = data.data() + ix * site_bytes;
      for (int k = 0; k < 2 * N; k++) {
        const double v = s[ix * 2 * N + k];
        if (prec == 64) put_double_be(p + k * w, v);
        else put_float_be(p + k * w, (float)v);
      }
      ck.accum(ix, p, site_bytes);
    }
    ok = ok && write_record(f, "scidac-binary-data", true, false, data.data(), data.size());
    const std::string x = checksum_xml(ck);
    ok = ok && write_record(f, "scidac-checksum", false, true, x.data(), x.size());
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
      set_error("write error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return CVCHIP_OK;
  });
}

// read_lime_contraction (reference contractions_io.c:430-505): the position-th
// scidac-binary-data record, precision from its length; the checksum of the
// data is returned and compared with the following scidac-checksum record
// when there is one.
extern "C" int cvchip_lime_read_contraction(const char *filename, double *s, int N, int position,
                                            const int dims[4], uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_read_contraction", [&]() -> int {
    if (!filename || !s || !dims) {
      set_error("NULL argument to cvchip_lime_read_contraction");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || N < 1 || position < 0) {
      set_error("invalid contraction read request (N %d, position %d)", N, position);
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "rb");
    if (!f) {
      set_error("could not open %s", filename);
      return CVCHIP_ERR_ARG;
    }
    std::vector<Record> recs;
    if (!scan(f, recs)) {
      fclose(f);
      set_error("%s is not a LIME file", filename);
      return CVCHIP_ERR_ARG;
    }
    int got = -1;
    size_t k = 0;
    for (; k < recs.size(); k++)
      if (recs[k].type == "scidac-binary-data" && ++got == position) break;
    if (k == recs.size()) {
      fclose(f);
      set_error("no scidac-binary-data record %d in %s", position, filename);
      return CVCHIP_ERR_ARG;
    }
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    int prec;
    if (recs[k].bytes == V * 2 * N * 8) prec = 64;
    else if (recs[k].bytes == V * 2 * N * 4) prec = 32;
    else {
      fclose(f);
      set_error("wrong length %llu of the contraction record in %s", (unsigned long long)recs[k].bytes, filename);
      return CVCHIP_ERR_ARG;
    }
    std::vector<unsigned char> data;
    if (!read_payload(f, recs[k], data)) {
      fclose(f);
      set_error("read error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    const size_t w = prec == 64 ? 8 : 4, site_bytes = (size_t)2 * N * w;
    Checksum ck;
    for (uint64_t ix = 0; ix < V; ix++) {
      const unsigned char *p = data.data() + ix * site_bytes;
      ck.accum(ix, p, site_bytes);
      for (int j = 0; j < 2 * N; j++)
        s[ix * 2 * N + j] = prec == 64 ? get_double_be(p + j * w) : (double)get_float_be(p + j * w);
    }
    int rc = CVCHIP_OK;
    if (k + 1 < recs.size() && recs[k + 1].type == "scidac-checksum") {
      std::vector<unsigned char> xml;
      uint32_t a, b;
      if (read_payload(f, recs[k + 1], xml) && parse_checksum(xml, a, b) && (a != ck.a || b != ck.b)) {
        set_error("checksum mismatch in %s: file %08x %08x, data %08x %08x", filename, a, b, ck.a, ck.b);
        rc = CVCHIP_ERR_ARG;
      }
    }
    fclose(f);
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return rc;
  });
}

// ------------------------------------------------------------ gauge (ILDG)

// read_lime_gauge_field_doubleprec / _singleprec (reference io.c:202-347):
// the first ildg-binary-data record, sites in t, z, y, x order (x fastest),
// per site the links mu = x, y, z, t as 3x3 complex row-major, big-endian;
// into the reference layout gauge[V][4][18] (mu = t, x, y, z, sites t, x, y,
// z with z fastest).  The checksum rank is the file order.
extern "C" int cvchip_lime_read_gauge(const char *filename, const int dims[4], double *gauge,
                                      uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_read_gauge", [&]() -> int {
    if (!filename || !dims || !gauge) {
      set_error("NULL argument to cvchip_lime_read_gauge");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims)) {
      set_error("invalid lattice extents");
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "rb");
    if (!f) {
      set_error("could not open %s", filename);
      return CVCHIP_ERR_ARG;
    }
    std::vector<Record> recs;
    if (!scan(f, recs)) {
      fclose(f);
      set_error("%s is not a LIME file", filename);
      return CVCHIP_ERR_ARG;
    }
    size_t k = 0;
    while (k < recs.size() && recs[k].type != "ildg-binary-data") k++;
    if (k == recs.size()) {
      fclose(f);
      set_error("no ildg-binary-data record in %s", filename);
      return CVCHIP_ERR_ARG;
    }
    const int T = dims[0], LX = dims[1], LY = dims[2], LZ = dims[3];
    const uint64_t V = (uint64_t)T * LX * LY * LZ;
    int w;
    if (recs[k].bytes == V * 72 * 8) w = 8;
    else if (recs[k].bytes == V * 72 * 4) w = 4;
    else {
      fclose(f);
      set_error("wrong lattice size or precision in %s (%llu bytes)", filename, (unsigned long long)recs[k].bytes);
      return CVCHIP_ERR_ARG;
    }
    std::vector<unsigned char> data;
    if (!read_payload(f, recs[k], data)) {
      fclose(f);
      set_error("read error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    Checksum ck;
    uint64_t rank = 0;
    for (int t = 0; t < T; t++)
      for (int z = 0; z < LZ; z++)
        for (int y = 0; y < LY; y++)
          for (int x = 0; x < LX; x++, rank++) {
            const unsigned char *p = data.data() + rank * 72 * w;
            ck.accum(rank, p, (size_t)72 * w);
            const uint64_t ix = (((uint64_t)t * LX + x) * LY + y) * LZ + z;
            for (int m = 0; m < 4; m++) {
              const int mu = m == 3 ? 0 : m + 1;   // file order x, y, z, t
              for (int e = 0; e < 18; e++) {
                const unsigned char *q = p + (size_t)(m * 18 + e) * w;
                gauge[(ix * 4 + mu) * 18 + e] = w == 8 ? get_double_be(q) : (double)get_float_be(q);
              }
            }
          }
    int rc = CVCHIP_OK;
    for (size_t j = k + 1; j < recs.size(); j++)
      if (recs[j].type == "scidac-checksum") {
        std::vector<unsigned char> xml;
        uint32_t a, b;
        if (read_payload(f, recs[j], xml) && parse_checksum(xml, a, b) && (a != ck.a || b != ck.b)) {
          set_error("checksum mismatch in %s: file %08x %08x, data %08x %08x", filename, a, b, ck.a, ck.b);
          rc = CVCHIP_ERR_ARG;
        }
        break;
      }
    fclose(f);
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return rc;
  });
}

// An ILDG gauge file the reader above (and the reference's) accepts:
// "ildg-format" XML, "ildg-binary-data" (64 or 32 bit), "scidac-checksum".
extern "C" int cvchip_lime_write_gauge(const char *filename, const int dims[4], const double *gauge, int prec,
                                       uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_write_gauge", [&]() -> int {
    if (!filename || !dims || !gauge) {
      set_error("NULL argument to cvchip_lime_write_gauge");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || (prec != 64 && prec != 32)) {
      set_error("invalid gauge format (precision %d)", prec);
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "wb");
    if (!f) {
      set_error("could not open %s for writing", filename);
      return CVCHIP_ERR_ARG;
    }
    const int T = dims[0], LX = dims[1], LY = dims[2], LZ = dims[3];
    char msg[1024];
    snprintf(msg, sizeof(msg),
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<ildgFormat xmlns=\"http://www.lqcd.org/ildg\"\n"
             "            xmlns:xsi=\"http://www.w3.org/2001/XMLSchema-instance\"\n"
             "            xsi:schemaLocation=\"http://www.lqcd.org/ildg http://www.lqcd.org/ildg/filefmt.xsd\">\n"
             "  <version>1.0</version>\n  <field>su3gauge</field>\n  <precision>%d</precision>\n"
             "  <lx>%d</lx>\n  <ly>%d</ly>\n  <lz>%d</lz>\n  <lt>%d</lt>\n</ildgFormat>",
             prec, LX, LY, LZ, T);
    bool ok = write_record(f, "ildg-format", true, false, msg, strlen(msg));
    const uint64_t V = (uint64_t)T * LX * LY * LZ;
    const size_t w = prec == 64 ? 8 : 4;
    std::vector<unsigned char> data(V * 72 * w);
    Checksum ck;
    uint64_t rank = 0;
    for (int t = 0; t < T; t++)
      for (int z = 0; z < LZ; z++)
        for (int y = 0; y < LY; y++)
          for (int x = 0; x < LX; x++, rank++) {
            unsigned char *p = data.data() + rank * 72 * w;
            const uint64_t ix = (((uint64_t)t * LX + x) * LY + y) * LZ + z;
            for (int m = 0; m < 4; m++) {
              const int mu = m == 3 ? 0 : m + 1;
              for (int e = 0; e < 18; e++) {
                const double v = gauge[(ix * 4 + mu) * 18 + e];
                if (w == 8) put_double_be(p + (size_t)(m * 18 + e) * w, v);
                else put_float_be(p + (size_t)(m * 18 + e) * w, (float)v);
              }
            }
            ck.accum(rank, p, (size_t)72 * w);
          }
    ok = ok && write_record(f, "ildg-binary-data", false, false, data.data(), data.size());
    const std::string x = checksum_xml(ck);
    ok = ok && write_record(f, "scidac-checksum", false, true, x.data(), x.size());
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
      set_error("write error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return CVCHIP_OK;
  });
}

// ------------------------------------------------------------ spinors (propagators)

namespace {

// a spinor field of the reference layout s[24 ix + 2 (3 spin + colour) + re/im]
// to / from the file's sites in t, z, y, x order (x fastest), 24 big-endian
// reals per site (read_binary_spinor_data / write_binary_spinor_data,
// reference propagator_io.c:141-200 and 344-405); rank = file position
void spinor_to_file(const double *s, const int d[4], int w, std::vector<unsigned char> &data, Checksum &ck) {
  const int T = d[0], LX = d[1], LY = d[2], LZ = d[3];
  data.resize((size_t)T * LX * LY * LZ * 24 * w);
  uint64_t rank = 0;
  for (int t = 0; t < T; t++)
    for (int z = 0; z < LZ; z++)
      for (int y = 0; y < LY; y++)
        for (int x = 0; x < LX; x++, rank++) {
          unsigned char *p = data.data() + rank * 24 * w;
          const uint64_t ix = (((uint64_t)t * LX + x) * LY + y) * LZ + z;
          for (int e = 0; e < 24; e++) {
            if (w == 8) put_double_be(p + e * 8, s[ix * 24 + e]);
            else put_float_be(p + e * 4, (float)s[ix * 24 + e]);
          }
          ck.accum(rank, p, (size_t)24 * w);
        }
}

void spinor_from_file(double *s, const int d[4], int w, const std::vector<unsigned char> &data, Checksum &ck) {
  const int T = d[0], LX = d[1], LY = d[2], LZ = d[3];
  uint64_t rank = 0;
  for (int t = 0; t < T; t++)
    for (int z = 0; z < LZ; z++)
      for (int y = 0; y < LY; y++)
        for (int x = 0; x < LX; x++, rank++) {
          const unsigned char *p = data.data() + rank * 24 * w;
          ck.accum(rank, p, (size_t)24 * w);
          const uint64_t ix = (((uint64_t)t * LX + x) * LY + y) * LZ + z;
          for (int e = 0; e < 24; e++) s[ix * 24 + e] = w == 8 ? get_double_be(p + e * 8) : (double)get_float_be(p + e * 4);
        }
}

// read_cmi (reference propagator_io.c:842-892): raw host-order floats, sites in
// x, y, z, t order (t fastest), 24 per site, scaled by 2 kappa
int read_cmi(FILE *f, const char *filename, const int d[4], double kappa, double *s) {
  const int T = d[0], LX = d[1], LY = d[2], LZ = d[3];
  const uint64_t V = (uint64_t)T * LX * LY * LZ;
  if (fseek(f, 0, SEEK_END) != 0 || (uint64_t)ftell(f) != V * 24 * sizeof(float)) {
    set_error("%s is neither a LIME file with a scidac-binary-data record nor a CMI file of this volume", filename);
    return CVCHIP_ERR_ARG;
  }
  std::vector<float> buf(V * 24);
  if (fseek(f, 0, SEEK_SET) != 0 || fread(buf.data(), sizeof(float), buf.size(), f) != buf.size()) {
    set_error("read error on %s", filename);
    return CVCHIP_ERR_ARG;
  }
  const double k2 = 2.0 * kappa;
  uint64_t pos = 0;
  for (int x = 0; x < LX; x++)
    for (int y = 0; y < LY; y++)
      for (int z = 0; z < LZ; z++)
        for (int t = 0; t < T; t++, pos++) {
          const uint64_t ix = (((uint64_t)t * LX + x) * LY + y) * LZ + z;
          for (int e = 0; e < 24; e++) s[ix * 24 + e] = k2 * (double)buf[pos * 24 + e];
        }
  return CVCHIP_OK;
}

// write_checksum (reference propagator_io.c:410-466): its own message, XML on one line
std::string spinor_checksum_xml(const Checksum &c) {
  char m[512];
  snprintf(m, sizeof(m),
           "<?xml version=\"1.0\" encoding=\"UTF-8\"?><scidacChecksum><version>1.0</version><suma>%#010x</suma>"
           "<sumb>%#010x</sumb></scidacChecksum>",
           c.a, c.b);
  return m;
}

}  // namespace

// read_lime_spinor (reference propagator_io.c:771-840): the position-th
// scidac-binary-data record (0: up, 1: down of a two-flavour file), 64 or 32
// bit by its length; a file without such a record is read as CMI (kappa is
// used only there).  The data checksum is returned and compared with a
// scidac-checksum record that follows the data record.
extern "C" int cvchip_lime_read_spinor(const char *filename, const int dims[4], int position, double kappa, double *s,
                                       uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_read_spinor", [&]() -> int {
    if (!filename || !dims || !s) {
      set_error("NULL argument to cvchip_lime_read_spinor");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || position < 0) {
      set_error("invalid spinor read request (position %d)", position);
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "rb");
    if (!f) {
      set_error("could not open %s", filename);
      return CVCHIP_ERR_ARG;
    }
    std::vector<Record> recs;
    size_t k = 0;
    const bool lime = scan(f, recs);
    if (lime) {
      int got = -1;
      for (; k < recs.size(); k++)
        if (recs[k].type == "scidac-binary-data" && ++got == position) break;
    }
    if (!lime || k == recs.size()) {
      const int rc = read_cmi(f, filename, dims, kappa, s);
      fclose(f);
      if (rc == CVCHIP_OK && checksum) checksum[0] = checksum[1] = 0;
      return rc;
    }
    const uint64_t V = (uint64_t)dims[0] * dims[1] * dims[2] * dims[3];
    int w;
    if (recs[k].bytes == V * 24 * 8) w = 8;
    else if (recs[k].bytes == V * 24 * 4) w = 4;
    else {
      fclose(f);
      set_error("wrong length %llu of the spinor record in %s (expected %llu)", (unsigned long long)recs[k].bytes,
                filename, (unsigned long long)(V * 24 * 8));
      return CVCHIP_ERR_ARG;
    }
    std::vector<unsigned char> data;
    if (!read_payload(f, recs[k], data)) {
      fclose(f);
      set_error("read error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    Checksum ck;
    spinor_from_file(s, dims, w, data, ck);
    int rc = CVCHIP_OK;
    if (k + 1 < recs.size() && recs[k + 1].type == "scidac-checksum") {
      std::vector<unsigned char> xml;
      uint32_t a, b;
      if (read_payload(f, recs[k + 1], xml) && parse_checksum(xml, a, b) && (a != ck.a || b != ck.b)) {
        set_error("checksum mismatch in %s: file %08x %08x, data %08x %08x", filename, a, b, ck.a, ck.b);
        rc = CVCHIP_ERR_ARG;
      }
    }
    fclose(f);
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return rc;
  });
}

// write_lime_spinor (reference propagator_io.c:582-647): a "scidac-binary-data"
// record (64 or 32 bit) and a "scidac-checksum" record, appended to the file
// or in a new one.
extern "C" int cvchip_lime_write_spinor(const char *filename, const int dims[4], const double *s, int append, int prec,
                                        uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_write_spinor", [&]() -> int {
    if (!filename || !dims || !s) {
      set_error("NULL argument to cvchip_lime_write_spinor");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || (prec != 64 && prec != 32)) {
      set_error("invalid spinor format (precision %d)", prec);
      return CVCHIP_ERR_ARG;
    }
    std::vector<unsigned char> data;
    Checksum ck;
    spinor_to_file(s, dims, prec == 64 ? 8 : 4, data, ck);
    FILE *f = fopen(filename, append ? "ab" : "wb");
    if (!f) {
      set_error("could not open %s for writing", filename);
      return CVCHIP_ERR_ARG;
    }
    bool ok = write_record(f, "scidac-binary-data", false, true, data.data(), data.size());
    const std::string x = spinor_checksum_xml(ck);
    ok = ok && write_record(f, "scidac-checksum", true, false, x.data(), x.size());
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
      set_error("write error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    if (checksum) {
      checksum[0] = ck.a;
      checksum[1] = ck.b;
    }
    return CVCHIP_OK;
  });
}

// write_propagator (reference propagator_io.c:125-133, the LIME build): a new
// file with the "etmc-propagator-format" record (write_propagator_format,
// propagator_io.c:522-576, one flavour), then the spinor as write_lime_spinor.
// A second flavour is appended with cvchip_lime_write_spinor(append = 1).
extern "C" int cvchip_lime_write_propagator(const char *filename, const int dims[4], const double *s, int prec,
                                            uint32_t checksum[2]) {
  return cvchip::guarded("cvchip_lime_write_propagator", [&]() -> int {
    if (!filename || !dims || !s) {
      set_error("NULL argument to cvchip_lime_write_propagator");
      return CVCHIP_ERR_NULL;
    }
    if (!check_dims(dims) || (prec != 64 && prec != 32)) {
      set_error("invalid spinor format (precision %d)", prec);
      return CVCHIP_ERR_ARG;
    }
    FILE *f = fopen(filename, "wb");
    if (!f) {
      set_error("could not open %s for writing", filename);
      return CVCHIP_ERR_ARG;
    }
    char msg[512];
    snprintf(msg, sizeof(msg),
             "<?xml version=\"1.0\" encoding=\"UTF-8\"?>\n<etmcFormat>\n<field>diracFermion</field>\n<precision>%d"
             "</precision>\n<flavours>%d</flavours>\n<lx>%d</lx>\n<ly>%d</ly>\n<lz>%d</lz>\n<lt>%d</lt>\n</etmcFormat>",
             prec, 1, dims[1], dims[2], dims[3], dims[0]);
    bool ok = write_record(f, "etmc-propagator-format", true, false, msg, strlen(msg));
    ok = (fclose(f) == 0) && ok;
    if (!ok) {
      set_error("write error on %s", filename);
      return CVCHIP_ERR_ARG;
    }
    return cvchip_lime_write_spinor(filename, dims, s, 1, prec, checksum);
  });
}
