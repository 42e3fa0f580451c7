// Shared pieces of the command-line drivers: option parser, stage timer,
// parallel readers of the atoms and LCAO text files, and engine checks.
#pragma once

#include <sys/time.h>

#include <algorithm>
#include <charconv>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <map>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "clat.h"

namespace cli {

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

[[noreturn]] inline void failf(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  std::vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Fail(buf);
}

inline void check(int rc) {
  if (rc != CLAT_OK) throw Fail(clat_last_error());
}

// ---------------------------------------------------------------------------
// Options: --name=value, --flag, short flags, positional arguments; unknown
// names are errors (src/Parser.cpp:35-99).
// ---------------------------------------------------------------------------
class Options {
 public:
  Options& value(const std::string& n) { values_ok_.insert(n); return *this; }
  Options& flag(const std::string& n) { flags_ok_.insert(n); return *this; }

  void parse(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
      std::string s = argv[i];
      if (s.size() > 1 && s[0] == '-') {
        if (s[1] == '-') {
          auto eq = s.find('=', 2);
          if (eq != std::string::npos) {
            std::string name = s.substr(2, eq - 2);
            if (values_ok_.count(name)) values_[name] = s.substr(eq + 1);
            else if (flags_ok_.count(name)) failf("flag '%s' does not need a value", name.c_str());
            else failf("unknown value '%s'", name.c_str());
          } else {
            std::string name = s.substr(2);
            if (flags_ok_.count(name)) flags_.insert(name);
            else if (values_ok_.count(name)) failf("missing value for parameter '%s'", name.c_str());
            else failf("unknown flag '%s'", name.c_str());
          }
        } else {
          failf("unknown flag '%c'", s[1]);
        }
      } else {
        args_.push_back(s);
      }
    }
  }

  bool has(const std::string& n, std::string& v) const {
    auto it = values_.find(n);
    if (it == values_.end()) return false;
    v = it->second;
    return true;
  }
  bool flag_set(const std::string& n) const { return flags_.count(n) != 0; }
  const std::string& get(const std::string& n) const {
    auto it = values_.find(n);
    if (it == values_.end()) failf("parameter '%s' not specified", n.c_str());
    return it->second;
  }
  const std::vector<std::string>& args() const { return args_; }

 private:
  std::set<std::string> values_ok_, flags_ok_, flags_;
  std::map<std::string, std::string> values_;
  std::vector<std::string> args_;
};

// ---------------------------------------------------------------------------
// Stage timer printing "<stage>... (x.xxx s)" on stderr (src/Timer.cpp:18-35).
// ---------------------------------------------------------------------------
class Timer {
 public:
  ~Timer() { stop(); }
  void start(const char* msg) {
    stop();
    std::fprintf(stderr, "%s... ", msg);
    std::fflush(stderr);
    gettimeofday(&t0_, nullptr);
    active_ = true;
  }
  void stop() {
    if (!active_) return;
    struct timeval now;
    gettimeofday(&now, nullptr);
    const double t = (now.tv_usec - t0_.tv_usec) * 1.0e-6 + (double)(now.tv_sec - t0_.tv_sec);
    std::fprintf(stderr, "(%.3f s)\n", t);
    std::fflush(stderr);
    active_ = false;
  }

 private:
  struct timeval t0_ {};
  bool active_ = false;
};

// ---------------------------------------------------------------------------
// Text input.  Whole-file read, line index, parallel number parsing with
// std::from_chars (correctly rounded, like sscanf's %lf) and a strtod
// fallback for the forms from_chars does not take (hex floats).
// ---------------------------------------------------------------------------
struct TextFile {
  std::string data;
  std::vector<size_t> starts;  // start offset of every line
  size_t line_end(size_t i) const {
    size_t e = (i + 1 < starts.size()) ? starts[i + 1] - 1 : data.size();
    return e;
  }
};

// line starts: 0 and every i + 1 with data[i] == '\n' and i + 1 < size,
// found in parallel slices with memchr and concatenated in order
inline TextFile read_text(const std::string& path, int threads = 1) {
  TextFile t;
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) failf("could not open file: %s", path.c_str());
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n > 0) {
    t.data.resize((size_t)n);
    if (std::fread(&t.data[0], 1, (size_t)n, f) != (size_t)n) {
      std::fclose(f);
      failf("could not read file: %s", path.c_str());
    }
  }
  std::fclose(f);
  const size_t size = t.data.size();
  if (size == 0) return t;
  const int T = std::max(1, std::min<int>(threads, (int)(size / (1 << 20)) + 1));
  std::vector<std::vector<size_t>> part((size_t)T);
  std::vector<std::thread> pool;
  const size_t chunk = (size + T - 1) / T;
  for (int k = 0; k < T; ++k) {
    pool.emplace_back([&, k] {
      const size_t a = std::min(size, (size_t)k * chunk), b = std::min(size, a + chunk);
      const char* base = t.data.data();
      const char* p = base + a;
      const char* e = base + b;
      while (p < e) {
        const char* q = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
        if (!q) break;
        const size_t i = (size_t)(q - base);
        if (i + 1 < size) part[(size_t)k].push_back(i + 1);
        p = q + 1;
      }
    });
  }
  for (auto& th : pool) th.join();
  size_t total = 1;
  for (const auto& v : part) total += v.size();
  t.starts.reserve(total);
  t.starts.push_back(0);
  for (const auto& v : part) t.starts.insert(t.starts.end(), v.begin(), v.end());
  return t;
}

inline bool is_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f' || c == '\n'; }

// parses one double at p (after optional whitespace); advances p
inline bool parse_double(const char*& p, const char* end, double& out) {
  while (p < end && is_space(*p)) ++p;
  if (p >= end) return false;
  const char* q = p;
  if (*q == '+') ++q;
  auto r = std::from_chars(q, end, out, std::chars_format::general);
  if (r.ec == std::errc() && !(r.ptr < end && (*r.ptr == 'x' || *r.ptr == 'X'))) {
    p = r.ptr;
    return true;
  }
  char buf[128];
  const size_t len = std::min<size_t>(sizeof buf - 1, (size_t)(end - p));
  std::memcpy(buf, p, len);
  buf[len] = 0;
  char* e = nullptr;
  out = std::strtod(buf, &e);
  if (e == buf) return false;
  p += (e - buf);
  return true;
}

inline int parse_threads() {
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return (int)std::min(16u, h);
}

template <typename F>
void parallel_for(size_t n, int threads, F&& f) {
  threads = std::max(1, std::min<int>(threads, (int)((n + 65535) / 65536)));
  if (threads == 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> pool;
  const size_t chunk = (n + threads - 1) / threads;
  for (int t = 0; t < threads; ++t) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    pool.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& th : pool) th.join();
}

// FunctionCollection::loadAtomsPositions (FunctionCollection.cpp:38-54):
// lines while three numbers parse; the first other line ends the list.
inline std::vector<double> read_atoms(const std::string& path, int threads) {
  TextFile t = read_text(path, threads);
  const size_t nl = t.starts.size();
  std::vector<double> xyz(3 * nl);
  std::vector<char> ok(nl, 0);
  parallel_for(nl, threads, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const char* p = t.data.data() + t.starts[i];
      const char* e = t.data.data() + t.line_end(i);
      ok[i] = parse_double(p, e, xyz[3 * i]) && parse_double(p, e, xyz[3 * i + 1]) &&
              parse_double(p, e, xyz[3 * i + 2]);
    }
  });
  size_t n = 0;
  while (n < nl && ok[n]) ++n;
  xyz.resize(3 * n);
  return xyz;
}

// FunctionCollection::loadFunctionFromFile (FunctionCollection.cpp:56-84):
// skip header lines, then n_atoms * n_orb lines of "re im" (extra columns
// ignored), atom-major / orbital-minor; missing or short lines are an error.
inline std::vector<double> read_lcao(const std::string& path, int skip, int64_t n_atoms, int n_orb,
                                     int threads) {
  TextFile t = read_text(path, threads);
  const size_t need = (size_t)n_atoms * n_orb;
  const size_t first = (size_t)skip;
  if (t.starts.size() < first + need) failf("file is truncated: %s", path.c_str());
  std::vector<double> c(2 * need);
  std::atomic<int> bad{0};
  parallel_for(need, threads, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const size_t li = first + i;
      const char* p = t.data.data() + t.starts[li];
      const char* e = t.data.data() + t.line_end(li);
      if (!parse_double(p, e, c[2 * i]) || !parse_double(p, e, c[2 * i + 1])) {
        bad.store(1);
        return;
      }
    }
  });
  if (bad.load()) failf("file is truncated: %s", path.c_str());
  return c;
}

// All LCAO files of a run, several at a time: up to `threads` files are read
// concurrently, the threads split between them.  Errors are reported for the
// first failing file in argument order, as a sequential read would.
inline std::vector<std::vector<double>> read_lcao_all(const std::vector<std::string>& paths, int skip,
                                                      int64_t n_atoms, int n_orb, int threads) {
  const int nf = (int)paths.size();
  std::vector<std::vector<double>> out((size_t)nf);
  if (nf == 0) return out;
  const int workers = std::max(1, std::min(threads, nf));
  const int per = std::max(1, threads / workers);
  std::vector<std::string> err((size_t)nf);
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int w = 0; w < workers; ++w) {
    pool.emplace_back([&] {
      for (int i = next.fetch_add(1); i < nf; i = next.fetch_add(1)) {
        try {
          out[(size_t)i] = read_lcao(paths[(size_t)i], skip, n_atoms, n_orb, per);
        } catch (const std::exception& ex) {
          err[(size_t)i] = ex.what();
        }
      }
    });
  }
  for (auto& th : pool) th.join();
  for (int i = 0; i < nf; ++i)
    if (!err[(size_t)i].empty()) throw Fail(err[(size_t)i]);
  return out;
}

inline std::string basename_of(const std::string& path) {
  auto s = path.find_last_of('/');
  return s == std::string::npos ? path : path.substr(s + 1);
}

}  // namespace cli
