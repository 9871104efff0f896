// SciDAC single-file output of propagators and fermions: the file QIO_WRITE_NAMED_OBJECT produces
// through QDPFileWriter (inline_qio_write_obj.cc:150-180, qio_write_obj_funcmap.cc:30-100) with the
// single-file volume format.  QIO, LIME and QDP++ are not in the reference tree; the layout below is
// the published SciDAC file format (USQCD QIO, format version 1.1) restated:
//
//  * LIME records: a 144-byte big-endian header (magic 0x456789ab, version 1, message-begin and
//    message-end bits in the top two bits of the next 16, the data length in 64 bits, a 128-byte type
//    string), then the data padded with zeros to a multiple of 8 bytes.
//  * message 1: "scidac-private-file-xml" (format version, spacetime, dims, volfmt),
//    "scidac-file-xml" (the caller's file XML);
//  * message 2: "scidac-private-record-xml" (date, recordtype 0 = lattice field, datatype,
//    precision, colors, spins, typesize, datacount), "scidac-record-xml" (the caller's record XML),
//    "scidac-binary-data", "scidac-checksum".
//  * binary data: sites in lexicographic order (x fastest), each site the QDP++ element in big-endian
//    IEEE floats: a propagator as [sink spin][source spin][sink colour][source colour][re, im]
//    (PSpinMatrix of PColorMatrix), a fermion as [spin][colour][re, im].
//  * checksum: per site the CRC-32 (IEEE 802.3 polynomial, zlib's crc32) of its bytes, rotated left by
//    (lexicographic index mod 29) into suma and by (index mod 31) into sumb, both combined by XOR.
//
// Host arrays use the library's host layout: sites checkerboard-major (parity * V/2 + lex / 2), then
// a propagator's sink spin*3+colour, then source spin*3+colour (complex doubles).
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <string>
#include <thread>
#include <vector>

#include "cmx_internal.h"

namespace cmx {

namespace {

uint32_t crc_table_entry(uint32_t c) {
  for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
  return c;
}
struct CrcTable {
  uint32_t t[256];
  CrcTable() {
    for (uint32_t i = 0; i < 256; ++i) t[i] = crc_table_entry(i);
  }
};
const CrcTable& crc_table() {
  static const CrcTable tab;
  return tab;
}

uint32_t rotl(uint32_t v, int r) { return r == 0 ? v : (v << r) | (v >> (32 - r)); }

void put_be16(unsigned char* p, uint16_t v) {
  p[0] = (unsigned char)(v >> 8);
  p[1] = (unsigned char)v;
}
void put_be32(unsigned char* p, uint32_t v) {
  for (int k = 0; k < 4; ++k) p[k] = (unsigned char)(v >> (24 - 8 * k));
}
void put_be64(unsigned char* p, uint64_t v) {
  for (int k = 0; k < 8; ++k) p[k] = (unsigned char)(v >> (56 - 8 * k));
}
uint64_t get_be(const unsigned char* p, int n) {
  uint64_t v = 0;
  for (int k = 0; k < n; ++k) v = (v << 8) | p[k];
  return v;
}

constexpr uint32_t LIME_MAGIC = 0x456789abu;
constexpr int LIME_HEADER = 144;

struct File {
  FILE* f = nullptr;
  ~File() {
    if (f) std::fclose(f);
  }
};

void write_record(FILE* f, const char* type, const void* data, uint64_t len, bool mb, bool me) {
  unsigned char h[LIME_HEADER];
  std::memset(h, 0, sizeof(h));
  put_be32(h, LIME_MAGIC);
  put_be16(h + 4, 1);
  put_be16(h + 6, (uint16_t)((mb ? 0x8000 : 0) | (me ? 0x4000 : 0)));
  put_be64(h + 8, len);
  std::strncpy(reinterpret_cast<char*>(h + 16), type, 127);
  CMX_REQUIRE(std::fwrite(h, 1, sizeof(h), f) == sizeof(h), CMX_ERR_ARG, "SciDAC write: I/O error");
  if (len) CMX_REQUIRE(std::fwrite(data, 1, len, f) == len, CMX_ERR_ARG, "SciDAC write: I/O error");
  static const unsigned char zeros[8] = {0};
  const size_t pad = (size_t)((8 - len % 8) % 8);
  if (pad) CMX_REQUIRE(std::fwrite(zeros, 1, pad, f) == pad, CMX_ERR_ARG, "SciDAC write: I/O error");
}

struct Record {
  std::string type;
  uint64_t len = 0;
  long offset = 0;  // of the data
  bool mb = false, me = false;
};

std::vector<Record> scan_records(FILE* f) {
  std::vector<Record> recs;
  unsigned char h[LIME_HEADER];
  long pos = 0;
  while (std::fread(h, 1, sizeof(h), f) == sizeof(h)) {
    CMX_REQUIRE(get_be(h, 4) == LIME_MAGIC, CMX_ERR_ARG, "SciDAC read: not a LIME record");
    Record r;
    const uint16_t flags = (uint16_t)get_be(h + 6, 2);
    r.mb = (flags & 0x8000) != 0;
    r.me = (flags & 0x4000) != 0;
    r.len = get_be(h + 8, 8);
    char t[129];
    std::memcpy(t, h + 16, 128);
    t[128] = 0;
    r.type = t;
    r.offset = pos + LIME_HEADER;
    recs.push_back(r);
    pos = r.offset + (long)((r.len + 7) / 8 * 8);
    CMX_REQUIRE(std::fseek(f, pos, SEEK_SET) == 0, CMX_ERR_ARG, "SciDAC read: truncated file");
  }
  return recs;
}

std::string read_record(FILE* f, const Record& r) {
  std::string s(r.len, '\0');
  CMX_REQUIRE(std::fseek(f, r.offset, SEEK_SET) == 0, CMX_ERR_ARG, "SciDAC read: truncated file");
  if (r.len) CMX_REQUIRE(std::fread(&s[0], 1, r.len, f) == r.len, CMX_ERR_ARG, "SciDAC read: truncated file");
  return s;
}

// text of the first <tag>...</tag> in xml ("" if absent)
std::string xml_value(const std::string& xml, const std::string& tag) {
  const std::string a = "<" + tag + ">", b = "</" + tag + ">";
  const size_t i = xml.find(a);
  if (i == std::string::npos) return "";
  const size_t j = xml.find(b, i + a.size());
  if (j == std::string::npos) return "";
  return xml.substr(i + a.size(), j - i - a.size());
}

struct Layout {
  int ncomp;       // complex numbers per site
  int bytes;       // per real
  long V, Vh;
  const char* datatype;
  // host complex index of file component k of the site at host index hs
  long host_offset(long hs, int k) const {
    if (ncomp == 144) {  // file [s1][s2][c1][c2] -> host [s1*3+c1][s2*3+c2]
      const int c2 = k % 3, c1 = (k / 3) % 3, s2 = (k / 9) % 4, s1 = k / 36;
      return hs * 144 + (s1 * 3 + c1) * 12 + (s2 * 3 + c2);
    }
    return hs * 12 + k;  // fermion: [s][c] in both
  }
};

Layout layout(const int dims[4], int kind, int prec) {
  CMX_REQUIRE(kind == 0 || kind == 1, CMX_ERR_ARG, "SciDAC: kind 0 (propagator) or 1 (fermion)");
  CMX_REQUIRE(prec == P64 || prec == P32, CMX_ERR_ARG, "SciDAC: precision DOUBLE or SINGLE");
  Layout L;
  L.V = 1;
  for (int m = 0; m < 4; ++m) {
    CMX_REQUIRE(dims[m] >= 2 && dims[m] % 2 == 0, CMX_ERR_ARG, "SciDAC: even lattice extents >= 2");
    L.V *= dims[m];
  }
  L.Vh = L.V / 2;
  L.ncomp = kind == 0 ? 144 : 12;
  L.bytes = prec == P64 ? 8 : 4;
  static const char* names[2][2] = {{"QDP_D3_DiracPropagator", "QDP_F3_DiracPropagator"},
                                    {"QDP_D3_DiracFermion", "QDP_F3_DiracFermion"}};
  L.datatype = names[kind][prec == P64 ? 0 : 1];
  return L;
}

long host_site(const int dims[4], long lex, long Vh) {
  long r = lex;
  int x[4];
  for (int m = 0; m < 4; ++m) {
    x[m] = (int)(r % dims[m]);
    r /= dims[m];
  }
  const int par = (x[0] + x[1] + x[2] + x[3]) & 1;
  return par * Vh + lex / 2;
}

// sites [lo, hi) of the binary data in file order, with their checksum contributions
void encode_sites(const int dims[4], const Layout& L, const double* host, long lo, long hi, unsigned char* out,
                  uint32_t& suma, uint32_t& sumb) {
  const CrcTable& ct = crc_table();
  const size_t site_bytes = (size_t)L.ncomp * 2 * L.bytes;
  for (long s = lo; s < hi; ++s) {
    unsigned char* p = out + (size_t)(s - lo) * site_bytes;
    const long hs = host_site(dims, s, L.Vh);
    for (int k = 0; k < L.ncomp; ++k) {
      const long o = L.host_offset(hs, k);
      for (int ri = 0; ri < 2; ++ri) {
        const double v = host[2 * o + ri];
        if (L.bytes == 8) {
          uint64_t u;
          std::memcpy(&u, &v, 8);
          put_be64(p, u);
        } else {
          const float fv = (float)v;
          uint32_t u;
          std::memcpy(&u, &fv, 4);
          put_be32(p, u);
        }
        p += L.bytes;
      }
    }
    uint32_t c = 0xFFFFFFFFu;
    const unsigned char* q = out + (size_t)(s - lo) * site_bytes;
    for (size_t b = 0; b < site_bytes; ++b) c = ct.t[(c ^ q[b]) & 0xFF] ^ (c >> 8);
    c ^= 0xFFFFFFFFu;
    suma ^= rotl(c, (int)(s % 29));
    sumb ^= rotl(c, (int)(s % 31));
  }
}

void decode_sites(const int dims[4], const Layout& L, const unsigned char* in, long lo, long hi, double* host,
                  uint32_t& suma, uint32_t& sumb) {
  const CrcTable& ct = crc_table();
  const size_t site_bytes = (size_t)L.ncomp * 2 * L.bytes;
  for (long s = lo; s < hi; ++s) {
    const unsigned char* p = in + (size_t)(s - lo) * site_bytes;
    uint32_t c = 0xFFFFFFFFu;
    for (size_t b = 0; b < site_bytes; ++b) c = ct.t[(c ^ p[b]) & 0xFF] ^ (c >> 8);
    c ^= 0xFFFFFFFFu;
    suma ^= rotl(c, (int)(s % 29));
    sumb ^= rotl(c, (int)(s % 31));
    const long hs = host_site(dims, s, L.Vh);
    for (int k = 0; k < L.ncomp; ++k) {
      const long o = L.host_offset(hs, k);
      for (int ri = 0; ri < 2; ++ri) {
        double v;
        if (L.bytes == 8) {
          const uint64_t u = get_be(p, 8);
          std::memcpy(&v, &u, 8);
        } else {
          const uint32_t u = (uint32_t)get_be(p, 4);
          float fv;
          std::memcpy(&fv, &u, 4);
          v = fv;
        }
        host[2 * o + ri] = v;
        p += L.bytes;
      }
    }
  }
}

int io_threads(long V) { return (int)std::max<long>(1, std::min<long>({16, (long)std::thread::hardware_concurrency(), V / 4096 + 1})); }

// runs f(lo, hi, suma, sumb) over chunks of sites on a few threads, combining the checksums
template <typename F> void over_chunks(long lo, long hi, int nt, uint32_t& suma, uint32_t& sumb, F&& f) {
  std::vector<uint32_t> a(nt, 0), b(nt, 0);
  std::vector<std::thread> th;
  const long per = (hi - lo + nt - 1) / nt;
  for (int t = 0; t < nt; ++t) {
    const long l = lo + t * per, h = std::min(hi, l + per);
    if (l >= h) continue;
    th.emplace_back([&, t, l, h]() { f(l, h, a[t], b[t]); });
  }
  for (auto& x : th) x.join();
  for (int t = 0; t < nt; ++t) suma ^= a[t], sumb ^= b[t];
}

}  // namespace

uint32_t scidac_crc32(const unsigned char* p, size_t n) {
  const CrcTable& ct = crc_table();
  uint32_t c = 0xFFFFFFFFu;
  for (size_t b = 0; b < n; ++b) c = ct.t[(c ^ p[b]) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

void scidac_write(const char* path, const int dims[4], int kind, int prec, const double* host, const char* file_xml,
                  const char* record_xml) {
  CMX_REQUIRE(path && host, CMX_ERR_ARG, "SciDAC write: null argument");
  const Layout L = layout(dims, kind, prec);
  File fh;
  fh.f = std::fopen(path, "wb");
  CMX_REQUIRE(fh.f, CMX_ERR_ARG, std::string("SciDAC write: cannot open ") + path);
  const std::string hdr = "<?xml version=\"1.0\" encoding=\"UTF-8\"?>";
  std::string pf = hdr + "<scidacFile><version>1.1</version><spacetime>4</spacetime><dims>";
  for (int m = 0; m < 4; ++m) pf += std::to_string(dims[m]) + " ";
  pf += "</dims><volfmt>0</volfmt></scidacFile>";
  const std::string fx = file_xml ? file_xml : "";
  write_record(fh.f, "scidac-private-file-xml", pf.data(), pf.size(), true, false);
  write_record(fh.f, "scidac-file-xml", fx.data(), fx.size(), false, true);
  char date[64];
  const std::time_t now = std::time(nullptr);
  std::strftime(date, sizeof(date), "%a %b %d %H:%M:%S %Y UTC", std::gmtime(&now));
  const size_t site_bytes = (size_t)L.ncomp * 2 * L.bytes;
  std::string pr = hdr + "<scidacRecord><version>1.1</version><date>" + date +
                   "</date><recordtype>0</recordtype><datatype>" + L.datatype + "</datatype><precision>" +
                   (L.bytes == 8 ? "D" : "F") + "</precision><colors>3</colors><spins>4</spins><typesize>" +
                   std::to_string(site_bytes) + "</typesize><datacount>1</datacount></scidacRecord>";
  const std::string rx = record_xml ? record_xml : "";
  write_record(fh.f, "scidac-private-record-xml", pr.data(), pr.size(), true, false);
  write_record(fh.f, "scidac-record-xml", rx.data(), rx.size(), false, false);
  // binary data: header, then the sites in slabs of about 64 MiB
  const uint64_t total = (uint64_t)L.V * site_bytes;
  {
    unsigned char h[LIME_HEADER];
    std::memset(h, 0, sizeof(h));
    put_be32(h, LIME_MAGIC);
    put_be16(h + 4, 1);
    put_be16(h + 6, 0);
    put_be64(h + 8, total);
    std::strncpy(reinterpret_cast<char*>(h + 16), "scidac-binary-data", 127);
    CMX_REQUIRE(std::fwrite(h, 1, sizeof(h), fh.f) == sizeof(h), CMX_ERR_ARG, "SciDAC write: I/O error");
  }
  uint32_t suma = 0, sumb = 0;
  const long slab = std::max<long>(1, (long)((64u << 20) / site_bytes));
  std::vector<unsigned char> buf((size_t)std::min(slab, L.V) * site_bytes);
  const int nt = io_threads(L.V);
  for (long lo = 0; lo < L.V; lo += slab) {
    const long hi = std::min(L.V, lo + slab);
    over_chunks(lo, hi, nt, suma, sumb, [&](long l, long h, uint32_t& a, uint32_t& b) {
      encode_sites(dims, L, host, l, h, buf.data() + (size_t)(l - lo) * site_bytes, a, b);
    });
    const size_t n = (size_t)(hi - lo) * site_bytes;
    CMX_REQUIRE(std::fwrite(buf.data(), 1, n, fh.f) == n, CMX_ERR_ARG, "SciDAC write: I/O error");
  }
  static const unsigned char zeros[8] = {0};
  const size_t pad = (size_t)((8 - total % 8) % 8);
  if (pad) CMX_REQUIRE(std::fwrite(zeros, 1, pad, fh.f) == pad, CMX_ERR_ARG, "SciDAC write: I/O error");
  char sums[160];
  std::snprintf(sums, sizeof(sums),
                "<scidacChecksum><version>1.0</version><suma>%x</suma><sumb>%x</sumb></scidacChecksum>", suma, sumb);
  const std::string ck = hdr + sums;
  write_record(fh.f, "scidac-checksum", ck.data(), ck.size(), false, true);
  CMX_REQUIRE(std::fflush(fh.f) == 0, CMX_ERR_ARG, "SciDAC write: I/O error");
}

void scidac_read(const char* path, const int dims[4], int kind, double* host, std::string* file_xml,
                 std::string* record_xml) {
  CMX_REQUIRE(path && host, CMX_ERR_ARG, "SciDAC read: null argument");
  File fh;
  fh.f = std::fopen(path, "rb");
  CMX_REQUIRE(fh.f, CMX_ERR_ARG, std::string("SciDAC read: cannot open ") + path);
  const std::vector<Record> recs = scan_records(fh.f);
  auto find = [&](const char* t) -> const Record* {
    for (const Record& r : recs)
      if (r.type == t) return &r;
    return nullptr;
  };
  const Record *rpf = find("scidac-private-file-xml"), *rpr = find("scidac-private-record-xml"),
               *rbin = find("scidac-binary-data"), *rck = find("scidac-checksum");
  CMX_REQUIRE(rpf && rpr && rbin, CMX_ERR_ARG, "SciDAC read: missing records");
  const std::string pf = read_record(fh.f, *rpf), pr = read_record(fh.f, *rpr);
  {
    std::string want;
    for (int m = 0; m < 4; ++m) want += std::to_string(dims[m]) + " ";
    std::string got = xml_value(pf, "dims");
    while (!got.empty() && got.back() == ' ') got.pop_back();
    while (!want.empty() && want.back() == ' ') want.pop_back();
    CMX_REQUIRE(got == want, CMX_ERR_ARG, "SciDAC read: lattice dims " + got + " differ from " + want);
  }
  const std::string prec = xml_value(pr, "precision");
  CMX_REQUIRE(prec == "D" || prec == "F", CMX_ERR_ARG, "SciDAC read: precision " + prec);
  const Layout L = layout(dims, kind, prec == "D" ? P64 : P32);
  CMX_REQUIRE(xml_value(pr, "datatype") == L.datatype, CMX_ERR_ARG,
              "SciDAC read: datatype " + xml_value(pr, "datatype") + ", expected " + L.datatype);
  const size_t site_bytes = (size_t)L.ncomp * 2 * L.bytes;
  CMX_REQUIRE(rbin->len == (uint64_t)L.V * site_bytes, CMX_ERR_ARG, "SciDAC read: binary data size");
  if (file_xml) {
    const Record* r = find("scidac-file-xml");
    *file_xml = r ? read_record(fh.f, *r) : "";
  }
  if (record_xml) {
    const Record* r = find("scidac-record-xml");
    *record_xml = r ? read_record(fh.f, *r) : "";
  }
  CMX_REQUIRE(std::fseek(fh.f, rbin->offset, SEEK_SET) == 0, CMX_ERR_ARG, "SciDAC read: truncated file");
  uint32_t suma = 0, sumb = 0;
  const long slab = std::max<long>(1, (long)((64u << 20) / site_bytes));
  std::vector<unsigned char> buf((size_t)std::min(slab, L.V) * site_bytes);
  const int nt = io_threads(L.V);
  for (long lo = 0; lo < L.V; lo += slab) {
    const long hi = std::min(L.V, lo + slab);
    const size_t n = (size_t)(hi - lo) * site_bytes;
    CMX_REQUIRE(std::fread(buf.data(), 1, n, fh.f) == n, CMX_ERR_ARG, "SciDAC read: truncated file");
    over_chunks(lo, hi, nt, suma, sumb, [&](long l, long h, uint32_t& a, uint32_t& b) {
      decode_sites(dims, L, buf.data() + (size_t)(l - lo) * site_bytes, l, h, host, a, b);
    });
  }
  if (rck) {
    const std::string ck = read_record(fh.f, *rck);
    const unsigned long a = std::strtoul(xml_value(ck, "suma").c_str(), nullptr, 16);
    const unsigned long b = std::strtoul(xml_value(ck, "sumb").c_str(), nullptr, 16);
    CMX_REQUIRE(a == suma && b == sumb, CMX_ERR_ARG, "SciDAC read: checksum mismatch");
  }
}

}  // namespace cmx
