// FCIDUMP reader and exact density-fitting factorisation (host C++; see ecg_fcidump.h).
#include "ecg_fcidump.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>
#include <zlib.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace ecg {
namespace {

[[noreturn]] void fail(const std::string& msg) { throw std::runtime_error(msg); }

// Line source over a plain or gzip-compressed file (zlib reads both).
struct Lines {
  gzFile f = nullptr;
  explicit Lines(const std::string& path) {
    f = gzopen(path.c_str(), "rb");
    if (!f) fail("cannot open FCIDUMP file " + path);
  }
  ~Lines() {
    if (f) gzclose(f);
  }
  bool next(std::string& line) {
    line.clear();
    char buf[4096];
    while (gzgets(f, buf, sizeof(buf))) {
      line += buf;
      if (!line.empty() && line.back() == '\n') break;
    }
    if (line.empty()) return false;
    while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
    return true;
  }
};

std::string strip(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::string strip_quotes(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && (s[a] == '"' || s[a] == '\'')) ++a;
  while (b > a && (s[b - 1] == '"' || s[b - 1] == '\'')) --b;
  return s.substr(a, b - a);
}

// Julia's tryparse(Int, s): optional sign, then decimal digits only
bool parse_int(const std::string& s, long& v) {
  size_t i = (!s.empty() && (s[0] == '+' || s[0] == '-')) ? 1 : 0;
  if (i >= s.size()) return false;
  for (size_t k = i; k < s.size(); ++k)
    if (!std::isdigit((unsigned char)s[k])) return false;
  errno = 0;
  v = std::strtol(s.c_str(), nullptr, 10);
  return errno == 0;
}

// decimal floating point (tryparse(Float64, s)); a Fortran D exponent is accepted too
bool parse_float(const std::string& s0, double& v) {
  if (s0.empty()) return false;
  std::string s = s0;
  for (char& c : s)
    if (c == 'D' || c == 'd') c = 'E';
  if (s.find('x') != std::string::npos || s.find('X') != std::string::npos) return false;
  char* end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}

std::vector<std::string> split_tokens(const std::string& s, bool comma) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == ' ' || c == '\t' || (comma && c == ',')) {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur += c;
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// read_header (dump.jl:360-405) with read_elements! (dump.jl:407-435): lines up to
// "/" or "&END" split at blanks and commas, '=' a token of its own; the token before
// each '=' is a (case-insensitive) key, the tokens up to the next key its values, typed
// by the first value (Int, else Float64, else String with quotes stripped).
void read_header(Lines& in, Fcidump& fd) {
  std::vector<std::string> tok;
  std::string line;
  bool ended = false;
  while (in.next(line)) {
    const std::string s = strip(line);
    if (s.empty()) continue;
    if (s == "/" || s == "&END") {
      ended = true;
      break;
    }
    std::string t;
    for (char c : s) {
      if (c == '=') t += " = ";
      else t += c;
    }
    for (auto& x : split_tokens(t, true)) tok.push_back(x);
  }
  if (!ended) fail("FCIDUMP header not terminated by '/' or '&END'");
  tok.push_back("\n");
  // read_elements!: values until the next '=' (whose preceding token is the next key)
  auto elements = [&](std::vector<std::string>* vals, char type, size_t ipos, std::string& key) {
    key.clear();
    std::string prev;
    while (ipos < tok.size()) {
      const std::string& el = tok[ipos];
      if (el == "=") {
        if (prev.empty()) fail("No variable name before '=' in the FCIDUMP header");
        key = prev;
        for (char& c : key) c = (char)std::toupper((unsigned char)c);
        break;
      }
      if (!prev.empty() && vals) {
        long iv;
        double fv;
        if (type == 'i' && !parse_int(prev, iv)) fail("FCIDUMP header: '" + prev + "' is not an integer");
        if (type == 'f' && !parse_float(prev, fv)) fail("FCIDUMP header: '" + prev + "' is not a number");
        vals->push_back(type == 's' ? strip_quotes(prev) : prev);
      }
      prev = el;
      ++ipos;
    }
    return ipos;
  };
  std::string key;
  size_t ipos = elements(nullptr, 's', 0, key);   // leading tokens ("&FCI") are comments
  while (ipos + 1 < tok.size()) {
    ++ipos;
    const std::string& first = tok[ipos];
    long iv;
    double fv;
    const char type = parse_int(first, iv) ? 'i' : (parse_float(first, fv) ? 'f' : 's');
    const std::string name = key;
    std::vector<std::string> vals{type == 's' ? strip_quotes(first) : first};
    ipos = elements(&vals, type, ipos + 1, key);
    fd.head[name] = vals;
    fd.head_type[name] = type;
  }
}

// NPY array of doubles in Fortran order (the form mnpymmap maps, src/tools/mnpy.jl:159-221);
// returns its shape and copies the data.
std::vector<double> read_npy(const std::string& path, std::vector<long>& shape) {
  const int fdesc = ::open(path.c_str(), O_RDONLY);
  if (fdesc < 0) fail("cannot open NPY file " + path);
  struct stat st;
  if (fstat(fdesc, &st) != 0) {
    ::close(fdesc);
    fail("cannot stat NPY file " + path);
  }
  const size_t size = (size_t)st.st_size;
  void* map = size ? mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fdesc, 0) : MAP_FAILED;
  ::close(fdesc);
  if (map == MAP_FAILED) fail("cannot map NPY file " + path);
  const unsigned char* b = static_cast<const unsigned char*>(map);
  auto bad = [&](const std::string& why) {
    munmap(map, size);
    fail("NPY file " + path + ": " + why);
  };
  if (size < 10 || std::memcmp(b, "\x93NUMPY", 6) != 0) bad("not a NPY file");
  const int major = b[6];
  size_t hlen, hoff;
  if (major == 1) {
    hlen = (size_t)b[8] | ((size_t)b[9] << 8);
    hoff = 10;
  } else if (major == 2 || major == 3) {
    if (size < 12) bad("truncated header");
    hlen = (size_t)b[8] | ((size_t)b[9] << 8) | ((size_t)b[10] << 16) | ((size_t)b[11] << 24);
    hoff = 12;
  } else {
    bad("unsupported NPY version");
  }
  if (hoff + hlen > size) bad("truncated header");
  const std::string hdr(reinterpret_cast<const char*>(b + hoff), hlen);
  auto field = [&](const std::string& k) {
    const size_t p = hdr.find("'" + k + "'");
    if (p == std::string::npos) bad("header lacks " + k);
    const size_t c = hdr.find(':', p);
    if (c == std::string::npos) bad("malformed header");
    return c + 1;
  };
  {
    size_t p = field("descr");
    const size_t q0 = hdr.find_first_of("'\"", p), q1 = hdr.find_first_of("'\"", q0 + 1);
    const std::string descr = hdr.substr(q0 + 1, q1 - q0 - 1);
    if (descr != "<f8" && descr != "=f8" && descr != "f8") bad("element type " + descr + " is not float64");
  }
  {
    size_t p = field("fortran_order");
    while (p < hdr.size() && hdr[p] == ' ') ++p;
    if (hdr.compare(p, 4, "True") != 0) bad("C-ordered array (the reader maps Fortran-ordered arrays)");
  }
  shape.clear();
  {
    size_t p = field("shape");
    const size_t a = hdr.find('(', p), e = hdr.find(')', a);
    if (a == std::string::npos || e == std::string::npos) bad("malformed shape");
    for (auto& t : split_tokens(hdr.substr(a + 1, e - a - 1), true)) {
      long v;
      if (!parse_int(t, v) || v < 0) bad("malformed shape");
      shape.push_back(v);
    }
  }
  size_t count = 1;
  for (long d : shape) count *= (size_t)d;
  const size_t doff = hoff + hlen;
  if (doff + 8 * count > size) bad("data shorter than its shape");
  std::vector<double> out(count);
  std::memcpy(out.data(), b + doff, 8 * count);
  munmap(map, size);
  return out;
}

bool file_exists(const std::string& p) {
  struct stat st;
  return ::stat(p.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

// read_integrals!(fd, dir) (dump.jl:442-461) for IUHF=0: int2 is <pq|rs> stored as
// [p, q, r + s(s+1)/2] for r <= s (0-based, an upper-triangular pair index of the last two
// indices), int1 is h[p,q]; the core energy is the header's ENUC.  Returns false when a
// named file is missing (the caller reads the text integrals instead).
bool read_npy_integrals(Fcidump& fd, const std::string& dir) {
  auto path_of = [&](const std::string& key) {
    const std::string f = fd.head.at(key)[0];
    return (!f.empty() && f[0] == '/') ? f : (dir.empty() ? f : dir + "/" + f);
  };
  if (!fd.has("NPY1")) fail("NPY1 option not found in fcidump");
  const std::string p2 = path_of("NPY2"), p1 = path_of("NPY1");
  for (const auto& p : {p2, p1})
    if (!file_exists(p)) {
      std::fprintf(stderr, "NPY-file %s not found. Continue with fcidump file.\n", p.c_str());
      return false;
    }
  if (!fd.has("ENUC")) fail("ENUC option not found in fcidump");
  const long n = fd.norb;
  std::vector<long> s2, s1;
  const std::vector<double> i2 = read_npy(p2, s2), i1 = read_npy(p1, s1);
  const long npair = n * (n + 1) / 2;
  if (s2.size() != 3 || s2[0] != n || s2[1] != n || s2[2] != npair)
    fail("NPY2 array must have shape (norb, norb, norb(norb+1)/2)");
  if (s1.size() != 2 || s1[0] != n || s1[1] != n) fail("NPY1 array must have shape (norb, norb)");
  fd.h = i1;
  fd.eri.assign((size_t)n * n * n * n, 0.0);
  for (long s = 0; s < n; ++s)
    for (long r = 0; r < n; ++r)
      for (long q = 0; q < n; ++q)
        for (long p = 0; p < n; ++p) {
          // <pq|rs> = <qp|sr> (particle symmetry) gives the pairs r > s
          const double v = r <= s ? i2[p + n * (q + n * (r + s * (s + 1) / 2))]
                                  : i2[q + n * (p + n * (s + r * (r + 1) / 2))];
          fd.eri[p + n * (r + n * (q + n * s))] = v;   // (pr|qs) = <pq|rs>
        }
  fd.int0 = fd.head_float("ENUC", 0.0);
  fd.from_npy = true;
  return true;
}

// read_integrals!(int1, int2, norb, fdfile, simtra) (dump.jl:573-604) with set_int1! and
// set_int2! (dump.jl:472-544): lines of five fields "value i j k l" (1-based; (ij|kl) for
// l > 0, h_ij for j > 0 = k = l, the core energy for i = j = k = l = 0, orbital energies
// i > 0 = j ignored); particle symmetry always, the full 8-fold symmetry unless ST > 0.
void read_text_integrals(Lines& in, Fcidump& fd) {
  const long n = fd.norb;
  const bool simtra = fd.simtra();
  fd.h.assign((size_t)n * n, 0.0);
  fd.eri.assign((size_t)n * n * n * n, 0.0);
  auto e = [&](long i, long j, long k, long l) -> double& { return fd.eri[i + n * (j + n * (k + n * l))]; };
  bool int0 = false;
  std::string line;
  while (in.next(line)) {
    const auto f = split_tokens(line, false);
    if (f.size() != 5) continue;
    double v;
    long idx[4];
    if (!parse_float(f[0], v)) fail("FCIDUMP: cannot parse the value of line: " + line);
    for (int k = 0; k < 4; ++k)
      if (!parse_int(f[k + 1], idx[k])) fail("FCIDUMP: cannot parse the indices of line: " + line);
    if (idx[0] > n || idx[1] > n || idx[2] > n || idx[3] > n) fail("Index larger than norb: " + line);
    const long i = idx[0] - 1, j = idx[1] - 1, k = idx[2] - 1, l = idx[3] - 1;
    if (l >= 0) {
      if (i < 0 || j < 0 || k < 0) fail("FCIDUMP: non-positive index in a two-electron line: " + line);
      e(i, j, k, l) = v;
      e(k, l, i, j) = v;
      if (!simtra) {
        e(j, i, k, l) = v; e(i, j, l, k) = v; e(j, i, l, k) = v;
        e(l, k, i, j) = v; e(k, l, j, i) = v; e(l, k, j, i) = v;
      }
    } else if (j >= 0) {
      if (i < 0) fail("FCIDUMP: non-positive index in a one-electron line: " + line);
      fd.h[i + n * j] = v;
      if (!simtra) fd.h[j + n * i] = v;
    } else if (i < 0) {
      fd.int0 = v;
      int0 = true;
    }
  }
  if (!int0) fail("No core energy found in fcidump. Incomplete file?");
}

}  // namespace

long Fcidump::head_int(const std::string& key, long def) const {
  auto it = head.find(key);
  if (it == head.end() || it->second.empty()) return def;
  long v;
  if (!parse_int(it->second[0], v)) fail("FCIDUMP header: " + key + " is not an integer");
  return v;
}

double Fcidump::head_float(const std::string& key, double def) const {
  auto it = head.find(key);
  if (it == head.end() || it->second.empty()) return def;
  double v;
  if (!parse_float(it->second[0], v)) fail("FCIDUMP header: " + key + " is not a number");
  return v;
}

Fcidump read_fcidump(const std::string& path) {
  Fcidump fd;
  Lines in(path);
  read_header(in, fd);
  fd.head.emplace("IUHF", std::vector<std::string>{"0"});   // defaults of read_header
  fd.head_type.emplace("IUHF", 'i');
  fd.head.emplace("ST", std::vector<std::string>{"0"});
  fd.head_type.emplace("ST", 'i');
  if (fd.head_int("IUHF", 0) > 0) fail("unrestricted FCIDUMP (IUHF > 0): the closed-shell path needs IUHF = 0");
  const long norb = fd.head_int("NORB", -1);
  if (norb < 0) fail("NORB option not found in fcidump");
  if (norb == 0 || norb > 4096) fail("FCIDUMP: NORB out of range");
  fd.norb = (int)norb;
  bool done = false;
  if (fd.has("NPY2")) {
    const size_t slash = path.find_last_of('/');
    done = read_npy_integrals(fd, slash == std::string::npos ? std::string() : path.substr(0, slash));
  }
  if (!done) read_text_integrals(in, fd);
  return fd;
}

FreezeCount freeze_orbs_in_dump(Fcidump& fd, const std::vector<int>& orbs) {
  const long n = fd.norb;
  const long nelec = fd.head_int("NELEC", -1);
  if (nelec < 0) fail("NELEC option not found in fcidump");
  if (nelec % 2 != 0 || fd.head_int("MS2", 0) != 0) fail("freeze_orbs_in_dump: the closed-shell path needs an even NELEC and MS2 = 0");
  const long nocc = nelec / 2;
  std::vector<char> frozen(n, 0);
  for (int o : orbs)
    if (o >= 1 && o <= n) frozen[o - 1] = 1;
  std::vector<long> core, kept;
  FreezeCount cnt;
  for (long p = 0; p < n; ++p) {
    if (!frozen[p]) {
      kept.push_back(p);
    } else if (p < nocc) {
      core.push_back(p);
      ++cnt.ncore;
    } else {
      ++cnt.nvirt;
    }
  }
  if (cnt.ncore == 0 && cnt.nvirt == 0) return cnt;
  auto eri = [&](long p, long q, long r, long s) { return fd.eri[p + n * (q + n * (r + n * s))]; };
  // core_fock[p,q] = sum_c 2 (pq|cc) - (pc|cq)
  std::vector<double> cf((size_t)n * n, 0.0);
  for (long q = 0; q < n; ++q)
    for (long p = 0; p < n; ++p) {
      double v = 0.0;
      for (long c : core) v += 2.0 * eri(p, q, c, c) - eri(p, c, c, q);
      cf[p + n * q] = v;
    }
  for (long c : core) fd.int0 += 2.0 * fd.h[c + n * c] + cf[c + n * c];
  const long m = (long)kept.size();
  std::vector<double> h((size_t)m * m), e((size_t)m * m * m * m);
  for (long b = 0; b < m; ++b)
    for (long a = 0; a < m; ++a) h[a + m * b] = fd.h[kept[a] + n * kept[b]] + cf[kept[a] + n * kept[b]];
  for (long d = 0; d < m; ++d)
    for (long c = 0; c < m; ++c)
      for (long b = 0; b < m; ++b)
        for (long a = 0; a < m; ++a) e[a + m * (b + m * (c + m * d))] = eri(kept[a], kept[b], kept[c], kept[d]);
  fd.h.swap(h);
  fd.eri.swap(e);
  fd.norb = (int)m;
  fd.head["NORB"] = {std::to_string(m)};
  fd.head_type["NORB"] = 'i';
  fd.head["NELEC"] = {std::to_string(nelec - 2 * cnt.ncore)};
  fd.head_type["NELEC"] = 'i';
  fd.head["ORBSYM"] = std::vector<std::string>(m, "1");
  fd.head_type["ORBSYM"] = 'i';
  return cnt;
}

void write_fcidump(const Fcidump& fd, const std::string& path, double tol) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) fail("cannot open " + path + " for writing");
  struct Closer {
    FILE* f;
    ~Closer() {
      if (f) std::fclose(f);
    }
  } closer{f};
  auto join = [&](const std::string& key) {
    const auto& v = fd.head.at(key);
    const bool quote = fd.head_type.count(key) && fd.head_type.at(key) == 's';
    std::string s;
    for (size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + (quote ? "\"" + v[i] + "\"" : v[i]);
    return s;
  };
  static const char* keys[] = {"NORB", "NELEC", "MS2", "ISYM", "ORBSYM"};
  std::fprintf(f, "&FCI\n");
  for (const char* k : keys)
    if (fd.has(k)) std::fprintf(f, " %s=%s,\n", k, join(k).c_str());
  for (char group : {'i', 'f', 's'})
    for (const auto& kv : fd.head) {
      const std::string& k = kv.first;
      const auto t = fd.head_type.find(k);
      if ((t == fd.head_type.end() ? 's' : t->second) != group) continue;
      if (std::find(std::begin(keys), std::end(keys), k) != std::end(keys)) continue;
      if ((k == "IUHF" || k == "ST" || k == "III") && !kv.second.empty() && std::strtod(kv.second[0].c_str(), nullptr) == 0.0)
        continue;
      std::fprintf(f, " %s=%s,\n", k.c_str(), join(k).c_str());
    }
  std::fprintf(f, "/\n");
  const long n = fd.norb;
  auto line = [&](double v, long i, long j, long k, long l) {
    std::fprintf(f, "%23.15e %3ld %3ld %3ld %3ld\n", v, i, j, k, l);
  };
  auto eri = [&](long p, long q, long r, long s) { return fd.eri[p + n * (q + n * (r + n * s))]; };
  // 1-based loops of the reference; the value of line "p q r s" is (pq|rs)
  if (fd.simtra()) {
    for (long p = 1; p <= n; ++p)
      for (long q = 1; q <= n; ++q) {
        for (long r = 1; r < p; ++r)
          for (long s = 1; s <= n; ++s) {
            const double v = eri(p - 1, q - 1, r - 1, s - 1);
            if (std::fabs(v) > tol) line(v, p, q, r, s);
          }
        for (long s = 1; s <= q; ++s) {
          const double v = eri(p - 1, q - 1, p - 1, s - 1);
          if (std::fabs(v) > tol) line(v, p, q, p, s);
        }
      }
    for (long p = 1; p <= n; ++p)
      for (long q = 1; q <= n; ++q) {
        const double v = fd.h[(p - 1) + n * (q - 1)];
        if (std::fabs(v) > tol) line(v, p, q, 0, 0);
      }
  } else {
    for (long p = 1; p <= n; ++p)
      for (long q = 1; q <= p; ++q)
        for (long r = 1; r <= p; ++r)
          for (long s = 1; s <= r; ++s) {
            if (r * (r - 1) / 2 + s > p * (p - 1) / 2 + q) continue;
            const double v = eri(p - 1, q - 1, r - 1, s - 1);
            if (std::fabs(v) > tol) line(v, p, q, r, s);
          }
    for (long p = 1; p <= n; ++p)
      for (long q = 1; q <= p; ++q) {
        const double v = fd.h[(p - 1) + n * (q - 1)];
        if (std::fabs(v) > tol) line(v, p, q, 0, 0);
      }
  }
  line(fd.int0, 0, 0, 0, 0);
  if (std::ferror(f)) fail("write error on " + path);
}

DfFactors df_factorise(const std::vector<double>& eri, int n, bool hermitian, double rtol) {
  const long N = (long)n * n;
  if ((long)eri.size() != N * N) fail("df_factorise: eri must hold norb^4 values");
  double amax = 0.0;
  for (double x : eri) amax = std::max(amax, std::fabs(x));
  for (long c = 0; c < N; ++c)
    for (long r = c + 1; r < N; ++r)
      if (std::fabs(eri[r + N * c] - eri[c + N * r]) > 1e-10 * std::max(1.0, amax))
        fail("integrals lack particle symmetry (pq|rs) = (rs|pq)");
  std::vector<double> R(eri);   // remaining matrix (right-looking update)
  double dmax = 0.0;
  for (long k = 0; k < N; ++k) dmax = std::max(dmax, hermitian ? R[k + N * k] : std::fabs(R[k + N * k]));
  std::vector<double> L;        // N x rank, column by column
  std::vector<double> D;
  std::vector<double> col(N);
  if (dmax > 0.0) {
    for (;;) {
      long piv = -1;
      double best = 0.0;
      for (long k = 0; k < N; ++k) {
        const double d = R[k + N * k];
        const double score = hermitian ? d : std::fabs(d);
        if (score > best) { best = score; piv = k; }
      }
      if (piv < 0 || best <= rtol * dmax) break;
      const double d = R[piv + N * piv];
      for (long r = 0; r < N; ++r) col[r] = R[r + N * piv] / d;
      col[piv] = 1.0;
      // R -= d col col^T (symmetric rank-one update)
      for (long c = 0; c < N; ++c) {
        const double dc = d * col[c];
        if (dc == 0.0) continue;
        double* Rc = &R[N * c];
        for (long r = 0; r < N; ++r) Rc[r] -= col[r] * dc;
      }
      for (long r = 0; r < N; ++r) R[r + N * piv] = R[piv + N * r] = 0.0;
      L.insert(L.end(), col.begin(), col.end());
      D.push_back(d);
      if ((long)D.size() >= N) break;
    }
  }
  if (hermitian) {
    double dmin = 0.0;
    for (long k = 0; k < N; ++k) dmin = std::min(dmin, R[k + N * k]);
    if (dmin < -1e-9 * dmax) fail("Hermitian integrals with a negative (pq|rs) eigenvalue");
  } else {
    double rmax = 0.0;
    for (double x : R) rmax = std::max(rmax, std::fabs(x));
    // a positive semidefinite remainder is bounded by its largest diagonal (<= rtol dmax);
    // an indefinite one with small diagonals is not
    if (rmax > std::max(100.0 * rtol * dmax, 1e-10 * std::max(1.0, amax)))
      fail("df_factorise: diagonal pivoting leaves a remainder (the integrals need 2x2 pivots)");
  }
  DfFactors out;
  out.naux = (int)D.size();
  bool neg = false;
  for (double d : D) neg = neg || d < 0.0;
  out.BL.resize((size_t)N * out.naux);
  if (neg) out.BR.resize((size_t)N * out.naux);
  for (int q = 0; q < out.naux; ++q) {
    const double s = std::sqrt(std::fabs(D[q]));
    for (long r = 0; r < N; ++r) {
      out.BL[r + N * q] = L[r + N * (long)q] * s;
      if (neg) out.BR[r + N * q] = L[r + N * (long)q] * (D[q] < 0.0 ? -s : s);
    }
  }
  return out;
}

}  // namespace ecg
