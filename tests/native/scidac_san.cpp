// Host-only check of the SciDAC writer and reader (chroma_amd/csrc/qio.cpp), built with AddressSanitizer
// and UndefinedBehaviorSanitizer by the Makefile beside it and run by tests/test_sanitizers.py.
//
//  * CRC-32 known answer: crc32("123456789") = 0xCBF43926 (the IEEE 802.3 check value).
//  * write -> read round trips of propagators and fermions in DOUBLE (exact) and SINGLE (float rounding)
//    on two lattices, with the file and record XML returned.
//  * the reader's error paths: wrong lattice dims, a file cut short inside the binary data, one flipped
//    byte in the binary data (checksum mismatch), a missing file.  Each must raise cmx::Error, not read
//    out of bounds.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "cmx_internal.h"

namespace cmx {
uint32_t scidac_crc32(const unsigned char* p, size_t n);
void scidac_write(const char* path, const int dims[4], int kind, int prec, const double* host, const char* file_xml,
                  const char* record_xml);
void scidac_read(const char* path, const int dims[4], int kind, double* host, std::string* file_xml,
                 std::string* record_xml);
}  // namespace cmx

static int failures = 0;

static void check(bool ok, const std::string& what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
  if (!ok) ++failures;
}

template <class F>
static bool raises(F f) {
  try {
    f();
  } catch (const cmx::Error&) {
    return true;
  }
  return false;
}

static std::vector<unsigned char> slurp(const std::string& p) {
  std::vector<unsigned char> b;
  FILE* f = std::fopen(p.c_str(), "rb");
  if (!f) return b;
  std::fseek(f, 0, SEEK_END);
  b.resize((size_t)std::ftell(f));
  std::fseek(f, 0, SEEK_SET);
  if (std::fread(b.data(), 1, b.size(), f) != b.size()) b.clear();
  std::fclose(f);
  return b;
}

static void spit(const std::string& p, const unsigned char* d, size_t n) {
  FILE* f = std::fopen(p.c_str(), "wb");
  std::fwrite(d, 1, n, f);
  std::fclose(f);
}

int main(int argc, char** argv) {
  const std::string dir = argc > 1 ? argv[1] : "/tmp";
  const unsigned char kat[] = "123456789";
  check(cmx::scidac_crc32(kat, 9) == 0xCBF43926u, "crc32 check value");

  const int lattices[2][4] = {{4, 4, 2, 6}, {2, 2, 2, 2}};
  std::mt19937_64 rng(11);
  std::normal_distribution<double> nd;
  for (const auto& dims : lattices) {
    const long V = (long)dims[0] * dims[1] * dims[2] * dims[3];
    for (int kind = 0; kind < 2; ++kind) {
      const long n = V * (kind == 0 ? 144 : 12) * 2;
      std::vector<double> in((size_t)n), out((size_t)n);
      for (double& x : in) x = nd(rng);
      for (int prec : {(int)cmx::P64, (int)cmx::P32}) {
        const std::string path = dir + "/scidac_san_" + std::to_string(kind) + "_" + std::to_string(prec) + ".lime";
        const std::string tag = std::string(kind == 0 ? "propagator" : "fermion") + (prec == cmx::P64 ? " D " : " F ") +
                                std::to_string(dims[0]) + "x" + std::to_string(dims[1]) + "x" +
                                std::to_string(dims[2]) + "x" + std::to_string(dims[3]);
        cmx::scidac_write(path.c_str(), dims, kind, prec, in.data(), "<file>x</file>", "<rec>y</rec>");
        std::string fx, rx;
        std::fill(out.begin(), out.end(), 0.0);
        cmx::scidac_read(path.c_str(), dims, kind, out.data(), &fx, &rx);
        double err = 0.0;
        for (long i = 0; i < n; ++i) {
          const double want = prec == cmx::P64 ? in[(size_t)i] : (double)(float)in[(size_t)i];
          err = std::fmax(err, std::fabs(out[(size_t)i] - want));
        }
        check(err == 0.0 && fx == "<file>x</file>" && rx == "<rec>y</rec>", "round trip " + tag);

        const int wrong[4] = {dims[0], dims[1], dims[2], dims[3] + 2};
        check(raises([&] { cmx::scidac_read(path.c_str(), wrong, kind, out.data(), nullptr, nullptr); }),
              "wrong dims rejected " + tag);
        check(raises([&] { cmx::scidac_read(path.c_str(), dims, 1 - kind, out.data(), nullptr, nullptr); }),
              "wrong kind rejected " + tag);

        const std::vector<unsigned char> file = slurp(path);
        // the binary data starts after its header; the site data is the bulk of the file
        const size_t site_bytes = (size_t)(kind == 0 ? 144 : 12) * 2 * (prec == cmx::P64 ? 8 : 4);
        const size_t data_len = (size_t)V * site_bytes;
        size_t data_at = 0;
        for (size_t p = 0; p + 144 <= file.size(); p += 8)
          if (std::strncmp(reinterpret_cast<const char*>(&file[p + 16]), "scidac-binary-data", 18) == 0) {
            data_at = p + 144;
            break;
          }
        check(data_at > 0 && data_at + data_len < file.size(), "binary data located " + tag);
        const std::string cut = path + ".cut";
        spit(cut, file.data(), data_at + data_len / 2);
        check(raises([&] { cmx::scidac_read(cut.c_str(), dims, kind, out.data(), nullptr, nullptr); }),
              "truncated file rejected " + tag);
        std::vector<unsigned char> bad = file;
        bad[data_at + data_len / 3] ^= 0x10;
        const std::string flip = path + ".flip";
        spit(flip, bad.data(), bad.size());
        check(raises([&] { cmx::scidac_read(flip.c_str(), dims, kind, out.data(), nullptr, nullptr); }),
              "flipped byte rejected " + tag);
        std::remove(path.c_str());
        std::remove(cut.c_str());
        std::remove(flip.c_str());
      }
    }
  }
  check(raises([&] {
          std::vector<double> o(16 * 24);
          cmx::scidac_read((dir + "/does_not_exist.lime").c_str(), lattices[1], 1, o.data(), nullptr, nullptr);
        }),
        "missing file rejected");
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "PASSED", failures);
  return failures ? 1 : 0;
}
