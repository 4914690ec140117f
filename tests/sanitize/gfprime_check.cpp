// Driver of the pole-file reader and writer under ASan/UBSan: reads the two
// golden files (Nlat = 2, 4 elements), writes them back in both forms and
// reads those, and feeds the reader damaged and hostile files, each of
// which must fail with VCA_ERR_ARG (never crash or allocate what the file
// does not hold).  Usage: gfprime_check GOLDEN_DIR SCRATCH_DIR
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vca_internal.hpp"

namespace vca {
static char g_msg[512];
[[noreturn]] void fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(g_msg, sizeof g_msg, fmt, ap);
  va_end(ap);
  throw Error{code};
}
void set_error(const char *, ...) {}
}  // namespace vca

using vca::GfEntry;

static int failures = 0;
static void expect(bool ok, const char *what) {
  std::printf("%s %s\n", ok ? "ok  " : "FAIL", what);
  if (!ok) ++failures;
}

static int read_code(const std::string &path, bool fmt, size_t nelem, std::vector<GfEntry> *out) {
  try {
    auto el = vca::gfprime_read(path, fmt, nelem);
    if (out) *out = std::move(el);
    return 0;
  } catch (const vca::Error &e) {
    std::printf("     (%s)\n", vca::g_msg);
    return e.code;
  }
}

static void write_bytes(const std::string &path, const std::vector<char> &b) {
  FILE *f = std::fopen(path.c_str(), "wb");
  std::fwrite(b.data(), 1, b.size(), f);
  std::fclose(f);
}

static void put32(std::vector<char> &b, int32_t v) {
  const char *p = reinterpret_cast<const char *>(&v);
  b.insert(b.end(), p, p + 4);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const std::string gold = argv[1], tmp = argv[2];
  std::vector<GfEntry> a, b;
  expect(read_code(gold + "/gfprime_flang.restart", false, 4, &a) == 0, "golden unformatted");
  expect(read_code(gold + "/gfprime_gfortran_style.restart", true, 4, &b) == 0,
         "golden formatted");
  expect(a.size() == 4 && b.size() == 4, "four elements each");
  for (bool fmt : {false, true}) {
    const std::string p = tmp + (fmt ? "/rt.txt" : "/rt.bin");
    vca::gfprime_write(p, fmt, a);
    std::vector<GfEntry> c;
    expect(read_code(p, fmt, 4, &c) == 0, fmt ? "round trip formatted" : "round trip unformatted");
    bool same = c.size() == a.size();
    for (size_t k = 0; same && k < a.size(); ++k) {
      same = c[k].size() == a[k].size();
      for (size_t s = 0; same && s < a[k].size(); ++s) {
        same = c[k][s].size() == a[k][s].size();
        for (size_t q = 0; same && q < a[k][s].size(); ++q)
          same = c[k][s][q].p == a[k][s][q].p && c[k][s][q].w == a[k][s][q].w;
      }
    }
    expect(same, "round trip equal");
  }
  // damaged and hostile unformatted records
  {
    std::vector<char> r;
    put32(r, 12);  // marker: 12 bytes
    put32(r, 1);   // one state
    put32(r, 1);   // one channel
    put32(r, 1 << 28);  // 2^28 poles in a 12-byte record
    put32(r, 12);
    write_bytes(tmp + "/huge_count.bin", r);
    expect(read_code(tmp + "/huge_count.bin", false, 1, nullptr) == VCA_ERR_ARG,
           "pole count beyond the record");
    std::vector<char> t;
    put32(t, 8);
    put32(t, 1 << 27);  // states beyond the record
    put32(t, 1);
    put32(t, 8);
    write_bytes(tmp + "/huge_states.bin", t);
    expect(read_code(tmp + "/huge_states.bin", false, 1, nullptr) == VCA_ERR_ARG,
           "state count beyond the record");
    std::vector<char> u;
    put32(u, 400);  // marker promises 400 bytes, the file ends
    put32(u, 1);
    write_bytes(tmp + "/truncated.bin", u);
    expect(read_code(tmp + "/truncated.bin", false, 1, nullptr) == VCA_ERR_ARG,
           "truncated record");
    std::vector<char> v;
    put32(v, 4);
    put32(v, -5);  // negative count
    put32(v, 4);
    write_bytes(tmp + "/negative.bin", v);
    expect(read_code(tmp + "/negative.bin", false, 1, nullptr) == VCA_ERR_ARG, "negative count");
  }
  // hostile formatted text
  {
    const char *cases[][2] = {
        {" 1 1 268435456 (1.0,0.0)\n", "pole count beyond the text"},
        {" 134217728\n", "state count beyond the text"},
        {" 1 1 2 (1.0,0.0\n", "unterminated complex"},
        {" 1 1 1 (1.0,0.0) (x,0)\n", "not a number"},
        {" 1 1 1 0*(1.0,0.0)\n", "zero repeat"},
    };
    int n = 0;
    for (auto &c : cases) {
      const std::string p = tmp + "/bad" + std::to_string(n++) + ".txt";
      FILE *f = std::fopen(p.c_str(), "w");
      std::fputs(c[0], f);
      std::fclose(f);
      expect(read_code(p, true, 1, nullptr) == VCA_ERR_ARG, c[1]);
    }
  }
  std::printf("%s\n", failures ? "FAILED" : "ALL OK");
  return failures ? 1 : 0;
}
