// gfprime.cpp -- the cluster Green's function pole file (host code).
//
// The file holds one entry per element of impGmatrix, in the loop order of
// save_gfprime / read_gfprime (VCA_AUX_FUNX.f90:608-697): ilat, jlat,
// ispin, iorb, jorb with jorb innermost.  Each entry is what the GFmatrix
// defined-I/O procedures transfer (VCA_AUX_FUNX.f90:484-604):
//
//   Nstates; per state: Nchan; per channel: Nexc, poles(Nexc), weight(Nexc)
//
// with default integers (4 bytes) and complex(8) poles and weights.
//
// Unformatted: a Fortran sequential unformatted file.  The child transfers
// of a defined WRITE(UNFORMATTED) do not start records, so each entry is one
// record framed by 4-byte length markers (gfortran/flang default; records
// over 2 GiB are split into subrecords with negative markers, which the
// reader follows).
//
// Formatted: list-directed text, one record (line) per entry: integers,
// complex values as (re,im), and after each state the literal token \n
// that write_formatted emits as a character item (VCA_AUX_FUNX.f90:503).
// The reader is list-directed in spirit: it takes the numbers in order
// across line breaks, accepts r*value repeat counts and skips the \n
// tokens (which the reference's own read_formatted does not expect between
// states, so its files with more than one state read back only here).
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "vca_internal.hpp"

namespace vca {

namespace {

struct FileCloser {
  FILE *f;
  ~FileCloser() {
    if (f) std::fclose(f);
  }
};

// ---------------------------------------------------------- unformatted --
void put_i32(std::vector<char> &b, int32_t v) {
  const char *p = reinterpret_cast<const char *>(&v);
  b.insert(b.end(), p, p + 4);
}
void put_c128(std::vector<char> &b, std::complex<double> v) {
  const double d[2] = {v.real(), v.imag()};
  const char *p = reinterpret_cast<const char *>(d);
  b.insert(b.end(), p, p + 16);
}

void write_record(FILE *f, const std::vector<char> &b, const std::string &path) {
  if (b.size() > (size_t)INT32_MAX - 8)
    fail(VCA_ERR_ARG, "%s: GF entry of %zu bytes exceeds one Fortran record", path.c_str(),
         b.size());
  const int32_t n = (int32_t)b.size();
  if (std::fwrite(&n, 4, 1, f) != 1 || (n && std::fwrite(b.data(), 1, b.size(), f) != b.size()) ||
      std::fwrite(&n, 4, 1, f) != 1)
    fail(VCA_ERR_ARG, "%s: write failed (%s)", path.c_str(), std::strerror(errno));
}

bool read_record(FILE *f, std::vector<char> &b, const std::string &path) {
  b.clear();
  for (;;) {
    int32_t m1 = 0, m2 = 0;
    const size_t got = std::fread(&m1, 1, 4, f);
    if (got == 0 && b.empty()) return false;
    if (got != 4) fail(VCA_ERR_ARG, "%s: truncated record marker", path.c_str());
    const int64_t len = m1 < 0 ? -(int64_t)m1 : m1;
    const size_t o = b.size();
    b.resize(o + (size_t)len);
    if (len && std::fread(b.data() + o, 1, (size_t)len, f) != (size_t)len)
      fail(VCA_ERR_ARG, "%s: truncated record", path.c_str());
    if (std::fread(&m2, 1, 4, f) != 4 || (m2 != m1 && m2 != -m1))
      fail(VCA_ERR_ARG, "%s: record markers do not match (%d, %d)", path.c_str(), m1, m2);
    if (m1 >= 0) return true;  // last (or only) subrecord
  }
}

struct Cursor {
  const std::vector<char> &b;
  size_t pos = 0;
  const std::string &path;
  // n items of at least `bytes` bytes each must fit the rest of the record
  // (checked before anything is sized from a count read from the file)
  void need(int64_t n, size_t bytes) const {
    if ((uint64_t)n > (b.size() - pos) / bytes)
      fail(VCA_ERR_ARG, "%s: GF entry shorter than its counts", path.c_str());
  }
  int32_t i32() {
    if (pos + 4 > b.size()) fail(VCA_ERR_ARG, "%s: GF entry shorter than its counts", path.c_str());
    int32_t v;
    std::memcpy(&v, b.data() + pos, 4);
    pos += 4;
    return v;
  }
  std::complex<double> c128() {
    if (pos + 16 > b.size())
      fail(VCA_ERR_ARG, "%s: GF entry shorter than its counts", path.c_str());
    double d[2];
    std::memcpy(d, b.data() + pos, 16);
    pos += 16;
    return {d[0], d[1]};
  }
};

// ------------------------------------------------------------ formatted --
void fmt_c128(std::string &out, std::complex<double> v) {
  char buf[80];
  std::snprintf(buf, sizeof buf, " (%.17E,%.17E)", v.real(), v.imag());
  out += buf;
}
void fmt_i32(std::string &out, int32_t v) {
  char buf[24];
  std::snprintf(buf, sizeof buf, " %11d", v);
  out += buf;
}

// Sequential reader of list-directed values over the whole file.
struct TextIn {
  std::string s;
  size_t pos = 0;
  const std::string &path;
  int64_t repeat_left = 0;
  std::string repeat_tok;

  void skip_sep() {
    for (;;) {
      while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n' ||
                                s[pos] == '\r' || s[pos] == ','))
        ++pos;
      // the "\n" character item written after each state
      if (pos + 1 < s.size() && s[pos] == '\\' && s[pos + 1] == 'n' &&
          (pos + 2 == s.size() || std::strchr(" \t\r\n,", s[pos + 2]))) {
        pos += 2;
        continue;
      }
      return;
    }
  }
  std::string raw_token() {
    skip_sep();
    if (pos >= s.size()) fail(VCA_ERR_ARG, "%s: file ends before the last GF entry", path.c_str());
    size_t st = pos;
    // a repeat count r* in front of a value
    size_t q = pos;
    while (q < s.size() && s[q] >= '0' && s[q] <= '9') ++q;
    std::string rep;
    if (q > pos && q < s.size() && s[q] == '*') {
      rep = s.substr(pos, q - pos);
      pos = q + 1;
      st = pos;
    }
    std::string tok;
    if (pos < s.size() && s[pos] == '(') {
      const size_t e = s.find(')', pos);
      if (e == std::string::npos) fail(VCA_ERR_ARG, "%s: unterminated complex value", path.c_str());
      tok = s.substr(st, e + 1 - st);
      pos = e + 1;
    } else {
      while (pos < s.size() && !std::strchr(" \t\r\n,", s[pos])) ++pos;
      tok = s.substr(st, pos - st);
    }
    if (!rep.empty()) {
      const long r = std::strtol(rep.c_str(), nullptr, 10);
      if (r < 1) fail(VCA_ERR_ARG, "%s: bad repeat count %s", path.c_str(), rep.c_str());
      repeat_left = r - 1;
      repeat_tok = tok;
    }
    return tok;
  }
  std::string token() {
    if (repeat_left > 0) {
      --repeat_left;
      return repeat_tok;
    }
    return raw_token();
  }
  int32_t i32() {
    const std::string t = token();
    char *end = nullptr;
    const long v = std::strtol(t.c_str(), &end, 10);
    if (end == t.c_str() || *end) fail(VCA_ERR_ARG, "%s: expected an integer, read '%s'",
                                       path.c_str(), t.c_str());
    return (int32_t)v;
  }
  static double real_of(const std::string &t, const std::string &path) {
    std::string u = t;
    for (char &c : u)
      if (c == 'D' || c == 'd') c = 'E';  // Fortran double-precision exponents
    char *end = nullptr;
    const double v = std::strtod(u.c_str(), &end);
    while (end && *end == ' ') ++end;
    if (end == u.c_str() || (end && *end))
      fail(VCA_ERR_ARG, "%s: expected a real, read '%s'", path.c_str(), t.c_str());
    return v;
  }
  std::complex<double> c128() {
    const std::string t = token();
    if (t.size() < 5 || t.front() != '(' || t.back() != ')')
      fail(VCA_ERR_ARG, "%s: expected a complex (re,im), read '%s'", path.c_str(), t.c_str());
    const size_t c = t.find(',');
    if (c == std::string::npos)
      fail(VCA_ERR_ARG, "%s: expected a complex (re,im), read '%s'", path.c_str(), t.c_str());
    auto trim = [](std::string x) {
      const size_t a = x.find_first_not_of(" \t\r\n");
      const size_t b = x.find_last_not_of(" \t\r\n");
      return a == std::string::npos ? std::string() : x.substr(a, b - a + 1);
    };
    return {real_of(trim(t.substr(1, c - 1)), path),
            real_of(trim(t.substr(c + 1, t.size() - c - 2)), path)};
  }
  bool at_end() {
    if (repeat_left > 0) return false;
    skip_sep();
    return pos >= s.size();
  }
};

void check_count(int64_t n, const char *what, const std::string &path) {
  if (n < 0 || n > (int64_t)1 << 28)
    fail(VCA_ERR_ARG, "%s: bad %s count %" PRId64, path.c_str(), what, n);
}

}  // namespace

void gfprime_write(const std::string &path, bool formatted, const std::vector<GfEntry> &el) {
  FILE *f = std::fopen(path.c_str(), formatted ? "w" : "wb");
  if (!f) fail(VCA_ERR_ARG, "cannot open %s for writing (%s)", path.c_str(), std::strerror(errno));
  FileCloser fc{f};
  for (const GfEntry &e : el) {
    if (formatted) {
      std::string line;
      fmt_i32(line, (int32_t)e.size());
      for (const auto &st : e) {
        fmt_i32(line, (int32_t)st.size());
        for (const auto &ch : st) {
          fmt_i32(line, (int32_t)ch.p.size());
          for (auto v : ch.p) fmt_c128(line, v);
          for (auto v : ch.w) fmt_c128(line, v);
        }
        line += " \\n";
      }
      line += '\n';
      if (std::fwrite(line.data(), 1, line.size(), f) != line.size())
        fail(VCA_ERR_ARG, "%s: write failed (%s)", path.c_str(), std::strerror(errno));
    } else {
      std::vector<char> b;
      put_i32(b, (int32_t)e.size());
      for (const auto &st : e) {
        put_i32(b, (int32_t)st.size());
        for (const auto &ch : st) {
          put_i32(b, (int32_t)ch.p.size());
          for (auto v : ch.p) put_c128(b, v);
          for (auto v : ch.w) put_c128(b, v);
        }
      }
      write_record(f, b, path);
    }
  }
  if (std::fflush(f) != 0)
    fail(VCA_ERR_ARG, "%s: write failed (%s)", path.c_str(), std::strerror(errno));
}

std::vector<GfEntry> gfprime_read(const std::string &path, bool formatted, size_t nelem) {
  FILE *f = std::fopen(path.c_str(), formatted ? "r" : "rb");
  if (!f) fail(VCA_ERR_ARG, "cannot open %s (%s)", path.c_str(), std::strerror(errno));
  FileCloser fc{f};
  std::vector<GfEntry> el(nelem);
  if (formatted) {
    TextIn in{std::string(), 0, path};
    char buf[1 << 16];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) in.s.append(buf, n);
    // the vectors grow as values are read, so a count larger than the data
    // that follows it fails at the end of the text, not in an allocation
    for (size_t k = 0; k < nelem; ++k) {
      const int32_t ns = in.i32();
      check_count(ns, "state", path);
      for (int32_t a = 0; a < ns; ++a) {
        el[k].emplace_back();
        const int32_t nc = in.i32();
        check_count(nc, "channel", path);
        for (int32_t b = 0; b < nc; ++b) {
          el[k].back().emplace_back();
          GfChanIO &ch = el[k].back().back();
          const int32_t ne = in.i32();
          check_count(ne, "pole", path);
          for (int32_t e = 0; e < ne; ++e) ch.p.push_back(in.c128());
          for (int32_t e = 0; e < ne; ++e) ch.w.push_back(in.c128());
        }
      }
    }
    if (!in.at_end())
      fail(VCA_ERR_ARG, "%s: data after the last of the %zu GF entries", path.c_str(), nelem);
  } else {
    std::vector<char> rec;
    for (size_t k = 0; k < nelem; ++k) {
      if (!read_record(f, rec, path))
        fail(VCA_ERR_ARG, "%s: %zu GF entries, expected %zu", path.c_str(), k, nelem);
      Cursor c{rec, 0, path};
      const int32_t ns = c.i32();
      check_count(ns, "state", path);
      c.need(ns, 4);  // a state holds at least its channel count
      el[k].resize(ns);
      for (auto &st : el[k]) {
        const int32_t nc = c.i32();
        check_count(nc, "channel", path);
        c.need(nc, 4);  // a channel holds at least its pole count
        st.resize(nc);
        for (auto &ch : st) {
          const int32_t ne = c.i32();
          check_count(ne, "pole", path);
          c.need(ne, 32);  // a pole and its weight, complex(8) each
          ch.p.resize(ne);
          ch.w.resize(ne);
          for (auto &v : ch.p) v = c.c128();
          for (auto &v : ch.w) v = c.c128();
        }
      }
      if (c.pos != rec.size())
        fail(VCA_ERR_ARG, "%s: GF entry %zu holds %zu bytes beyond its counts", path.c_str(), k,
             rec.size() - c.pos);
    }
    if (read_record(f, rec, path))
      fail(VCA_ERR_ARG, "%s: more records than the %zu GF entries", path.c_str(), nelem);
  }
  return el;
}

}  // namespace vca
