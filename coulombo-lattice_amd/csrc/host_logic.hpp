// Host-side logic of the integral path: grid rounding, lattice detection,
// atom -> cell indexing, integral patterns, the vertex-cover planner and the
// export layout.  Header-only C++17, shared by the C-ABI library and the
// command-line drivers.  Errors are reported as clat::host::Error.
#pragma once

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace clat {
namespace host {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// ---------------------------------------------------------------------------
// Round::up<p1, p2, ...>(n, c)  (src/Round.hpp:20-27, 38-49): smallest
// integer >= n of the form c * p1^a1 * p2^a2 * ...
// ---------------------------------------------------------------------------
inline int64_t round_up_rec(int64_t n, int64_t c, const int* p, int np) {
  if (np == 1) {
    while (c < n) c *= p[0];
    return c;
  }
  int64_t best = round_up_rec(n, c, p + 1, np - 1);
  while (c < n) {
    c *= p[0];
    best = std::min(best, round_up_rec(n, c, p + 1, np - 1));
  }
  return best;
}

inline int round_up(int n, int c, const std::vector<int>& primes) {
  if (primes.empty() || c < 1) throw Error(1, "invalid rounding request");
  for (int p : primes)
    if (p <= 1) throw Error(1, "rounding factors must be > 1");
  const int64_t r = round_up_rec(n, c, primes.data(), (int)primes.size());
  if (r > std::numeric_limits<int>::max()) throw Error(1, "rounded size overflows int");
  return (int)r;
}

inline int round235(int n) { return round_up(n, 1, {2, 3, 5}); }
inline int round2357(int n) { return round_up(n, 1, {2, 3, 5, 7}); }

// ---------------------------------------------------------------------------
// Lattice detection (src/Broadcaster.cpp:8-43).  Floating-point gcd of the
// coordinate differences with the reference's 1e-10 / 1e-2 thresholds.
// ---------------------------------------------------------------------------
inline double float_gcd(double a, double b) {
  if (a < b) std::swap(a, b);
  while (true) {
    const double r = std::fmod(a, b);
    if (r < 1.0e-10) return b;
    if (r < 1.0e-2) return 0.0;
    a = b;
    b = r;
  }
}

struct AxisGrid {
  int size;
  double origin, step;
};

// values: one coordinate of every atom (copied, sorted here)
inline AxisGrid compute_axis(std::vector<double> v) {
  if (v.empty()) throw Error(1, "no atoms");
  std::sort(v.begin(), v.end());
  double step = NAN;
  for (size_t a = 1; a < v.size(); ++a) {
    const double next = v[a] - v[a - 1];
    if (next > 1.0e-10) {
      step = std::isnan(step) ? next : float_gcd(step, next);
      if (step == 0) throw Error(1, "could not find common grid step");
    }
  }
  AxisGrid g;
  g.origin = v[0];
  g.step = step;
  if (std::isnan(step)) {
    // every atom shares this coordinate: one grid plane.  The reference
    // evaluates lrint(0/NaN) here, which has no defined value.
    g.size = 1;
    g.step = 1.0;
  } else {
    const long s = std::lrint((v.back() - v.front()) / step) + 1;
    if (s <= 0 || s > std::numeric_limits<int>::max()) throw Error(1, "dimensions are too large");
    g.size = (int)s;
  }
  return g;
}

struct Lattice {
  int raw[3];
  int pad[3];
  double origin[3];
  double step[3];
};

inline Lattice detect_lattice(const double* xyz, int64_t n) {
  Lattice L;
  for (int d = 0; d < 3; ++d) {
    std::vector<double> v((size_t)n);
    for (int64_t a = 0; a < n; ++a) v[(size_t)a] = xyz[3 * a + d];
    AxisGrid g = compute_axis(std::move(v));
    L.raw[d] = g.size;
    L.origin[d] = g.origin;
    L.step[d] = g.step;
    L.pad[d] = round235(g.size);  // Broadcaster::round4FFT (Broadcaster.cpp:136-143)
  }
  const double cells = (double)L.pad[0] * L.pad[1] * L.pad[2];
  if (cells >= 2147483647.0) throw Error(1, "dimensions are too large");
  return L;
}

// Broadcaster.cpp:76-88 at one rank.
inline void cell_indices(const double* xyz, int64_t n, const int dim[3], const double origin[3],
                         const double step[3], int32_t* cell) {
  for (int64_t a = 0; a < n; ++a) {
    const long ix = std::lrint((xyz[3 * a + 0] - origin[0]) / step[0]);
    const long iy = std::lrint((xyz[3 * a + 1] - origin[1]) / step[1]);
    const long iz = std::lrint((xyz[3 * a + 2] - origin[2]) / step[2]);
    if (ix < 0 || iy < 0 || iz < 0 || ix >= dim[0] || iy >= dim[1] || iz >= dim[2])
      throw Error(1, "atom outside the grid");
    cell[a] = (int32_t)((iz * (long)dim[1] + iy) * dim[0] + ix);
  }
}

// ---------------------------------------------------------------------------
// Integral pattern list (src/Pattern.hpp).  Characters: '*', digits 1-9,
// 'h', 'e', any other letter a free variable bound on first use.  As in the
// reference only the first four characters are read; a shorter pattern reads
// the string terminator as a fourth free variable.
// ---------------------------------------------------------------------------
class Pattern {
 public:
  Pattern(const std::string& description, int hole_count) : holes_(hole_count) {
    std::string::size_type start = 0, comma;
    while ((comma = description.find(',', start)) != std::string::npos) {
      add(description.substr(start, comma - start));
      start = comma + 1;
    }
    add(description.substr(start));
  }

  bool match(int i0, int i1, int i2, int i3) const {
    for (const std::string& p : patterns_)
      if (match_one(p, i0, i1, i2, i3)) return true;
    return false;
  }

 private:
  int holes_;
  std::vector<std::string> patterns_;

  static void check_char(char c) {
    const bool ok = (c == '*') || (std::isalnum((unsigned char)c) && c != '0');
    if (!ok) {
      char msg[64];
      std::snprintf(msg, sizeof msg, "invalid character '%c' in pattern", c);
      throw Error(1, msg);
    }
  }

  void add(const std::string& p) {
    if (p.empty()) throw Error(1, "invalid empty pattern");
    for (char c : p) check_char(c);
    patterns_.push_back(p);
  }

  bool letter(char c, int i, int* assign) const {
    if (c == '*') return true;
    if (c >= '1' && c <= '9') return i == (c - '0');
    if (c == 'h') return i <= holes_;
    if (c == 'e') return i > holes_;
    const int k = (unsigned char)c & 127;
    if (assign[k]) return i == assign[k];
    assign[k] = i;
    return true;
  }

  bool match_one(const std::string& p, int i0, int i1, int i2, int i3) const {
    int assign[128] = {0};
    auto ch = [&](size_t k) { return k < p.size() ? p[k] : '\0'; };
    return letter(ch(0), i0, assign) && letter(ch(1), i1, assign) && letter(ch(2), i2, assign) &&
           letter(ch(3), i3, assign);
  }
};

// coulombo.cpp:135-148
inline std::vector<int32_t> enumerate_integrals(const Pattern& pat, int n) {
  std::vector<int32_t> q;
  for (int i1 = 1; i1 <= n; ++i1)
    for (int i2 = 1; i2 <= n; ++i2)
      for (int i3 = 1; i3 <= n; ++i3)
        for (int i4 = 1; i4 <= n; ++i4)
          if (pat.match(i1, i2, i3, i4)) {
            q.push_back(i1);
            q.push_back(i2);
            q.push_back(i3);
            q.push_back(i4);
          }
  return q;
}

// ---------------------------------------------------------------------------
// Greedy vertex cover (src/Graph.hpp:49-91).
// ---------------------------------------------------------------------------
// edges: sorted and unique (the iteration order of the reference's std::set)
inline std::vector<int> vertex_cover(int nv, const std::vector<std::pair<int, int>>& edges) {
  std::set<int> cover;
  for (const auto& e : edges)
    if (e.first == e.second) cover.insert(e.first);
  std::vector<std::set<int>> nb((size_t)nv);
  for (const auto& e : edges)
    if (e.first != e.second && !cover.count(e.first) && !cover.count(e.second)) {
      nb[e.first].insert(e.second);
      nb[e.second].insert(e.first);
    }
  auto cmp = [&nb](int v, int w) {
    const size_t vs = nb[v].size(), ws = nb[w].size();
    return (vs < ws) || (vs == ws && v < w);
  };
  std::set<int, decltype(cmp)> queue(cmp);
  for (int v = 0; v < nv; ++v)
    if (!nb[v].empty()) queue.insert(v);
  while (!queue.empty()) {
    int v = *queue.begin();
    if (nb[v].size() == 1) v = *nb[v].begin();
    else v = *queue.rbegin();
    queue.erase(v);
    cover.insert(v);
    for (int w : nb[v]) {
      if (queue.erase(w)) {
        nb[w].erase(v);
        if (!nb[w].empty()) queue.insert(w);
      }
    }
  }
  return std::vector<int>(cover.begin(), cover.end());
}

inline std::vector<int> vertex_cover(int nv, const std::set<std::pair<int, int>>& edges) {
  return vertex_cover(nv, std::vector<std::pair<int, int>>(edges.begin(), edges.end()));
}

// LSD radix sort of 64-bit keys, 16 bits per pass over the bits in use
inline void radix_sort_u64(std::vector<uint64_t>& v, int bits) {
  std::vector<uint64_t> tmp(v.size());
  for (int shift = 0; shift < bits; shift += 16) {
    std::vector<size_t> count(65537, 0);
    for (uint64_t x : v) ++count[((x >> shift) & 0xFFFF) + 1];
    for (int i = 0; i < 65536; ++i) count[i + 1] += count[i];
    for (uint64_t x : v) tmp[count[(x >> shift) & 0xFFFF]++] = x;
    v.swap(tmp);
  }
}

// ---------------------------------------------------------------------------
// MasterPlanner (src/Planner.hpp:65-145).
// ---------------------------------------------------------------------------
struct PlanProduct {
  int product;
  int conj;
};
struct PlanStep {
  int id;
  PlanProduct left, right;
};

// Planner.hpp:134-144 (1-based states; iL >= iR after the swap)
inline PlanProduct plan_product(int iL, int iR) {
  int conj = 0;
  if (iL < iR) {
    conj = 1;
    std::swap(iL, iR);
  }
  return PlanProduct{iL * (iL - 1) / 2 + iR - 1, conj};
}

// ordered = false: the steps in id order after the cover swap, for callers
// that need each step's products but not the execution order
// fn(k0, k1, t) over [0, n) in contiguous ranges, one per thread t: up to
// 16 host threads, one per 65536 items (the calling thread takes range 0)
template <typename F>
inline size_t parallel_ranges(size_t n, F&& fn) {
  const size_t nt = std::min<size_t>({(size_t)16, (size_t)std::max(1u, std::thread::hardware_concurrency()),
                                      std::max<size_t>(1, n / 65536)});
  if (nt <= 1) {
    fn((size_t)0, n, (size_t)0);
    return 1;
  }
  std::vector<std::thread> pool;
  for (size_t t = 1; t < nt; ++t) pool.emplace_back([&fn, n, nt, t] { fn(n * t / nt, n * (t + 1) / nt, t); });
  fn((size_t)0, n / nt, (size_t)0);
  for (auto& th : pool) th.join();
  return nt;
}

// The product graph's edges (left * P + right of every integral), sorted and
// unique: a bitmap over P x P when it is small (per-thread bitmaps OR-merged
// while they stay small), else a sort of the edge keys.
template <typename KeyOf>
inline std::vector<std::pair<int, int>> product_edges(size_t n, int64_t P, KeyOf key_of) {
  std::vector<std::pair<int, int>> edges;
  if (P * P <= ((int64_t)1 << 28)) {
    const size_t words = (size_t)((P * P + 63) / 64);
    std::vector<uint64_t> bits(words, 0);
    const size_t nt_max = words <= ((size_t)1 << 16) ? 16 : 1;
    std::vector<std::vector<uint64_t>> part(nt_max > 1 ? nt_max : 0);
    auto mark = [&](size_t k0, size_t k1, uint64_t* b) {
      for (size_t k = k0; k < k1; ++k) {
        const uint64_t e = key_of(k);
        b[e >> 6] |= (uint64_t)1 << (e & 63);
      }
    };
    if (nt_max > 1) {
      const size_t nt = parallel_ranges(n, [&](size_t k0, size_t k1, size_t t) {
        part[t].assign(words, 0);
        mark(k0, k1, part[t].data());
      });
      for (size_t t = 0; t < nt; ++t)
        for (size_t w = 0; w < words; ++w) bits[w] |= part[t][w];
    } else {
      mark(0, n, bits.data());
    }
    for (size_t w = 0; w < bits.size(); ++w)
      for (uint64_t b = bits[w]; b; b &= b - 1) {
        const uint64_t e = w * 64 + (uint64_t)__builtin_ctzll(b);
        edges.push_back({(int)(e / P), (int)(e % P)});
      }
  } else {
    std::vector<uint64_t> ek(n);
    for (size_t k = 0; k < n; ++k) ek[k] = key_of(k);
    std::sort(ek.begin(), ek.end());
    ek.erase(std::unique(ek.begin(), ek.end()), ek.end());
    for (uint64_t e : ek) edges.push_back({(int)(e / P), (int)(e % P)});
  }
  return edges;
}

// Every state index in 1..n_states (checked over all integrals in parallel).
inline void check_quads(const int32_t* quads, int64_t n, int n_states) {
  if (n > std::numeric_limits<int>::max()) throw Error(1, "too many integrals to compute");
  std::vector<char> bad(16, 0);  // a range check failed in thread t
  parallel_ranges((size_t)n, [&](size_t k0, size_t k1, size_t t) {
    for (size_t k = 4 * k0; k < 4 * k1; ++k)
      if (quads[k] < 1 || quads[k] > n_states) {
        bad[t] = 1;
        return;
      }
  });
  for (char b : bad)
    if (b) throw Error(1, "integral state index out of range");
}

// The vertex cover of the product graph (in_cover[p]) and the products that
// lead at least one step once the steps are swapped into the cover (leads[p]),
// from the integrals alone: what orders the prepares, ahead of building the
// steps themselves.
inline void plan_cover(const int32_t* quads, int64_t n, int n_states, std::vector<char>& in_cover,
                       std::vector<char>& leads) {
  if (n > std::numeric_limits<int>::max()) throw Error(1, "too many integrals to compute");
  const int64_t P = (int64_t)n_states * (n_states + 1) / 2;
  // the range check rides on the edge scan (one pass over the integrals); an
  // out-of-range integral sets the flag and contributes edge 0
  std::atomic<bool> bad{false};
  const auto edges = product_edges((size_t)n, P, [&](size_t k) -> uint64_t {
    const int32_t* q = quads + 4 * k;
    for (int j = 0; j < 4; ++j)
      if (q[j] < 1 || q[j] > n_states) {
        bad.store(true, std::memory_order_relaxed);
        return 0;
      }
    return (uint64_t)plan_product(q[0], q[3]).product * P + plan_product(q[1], q[2]).product;
  });
  if (bad.load()) throw Error(1, "integral state index out of range");
  in_cover.assign((size_t)P, 0);
  for (int v : vertex_cover((int)P, edges)) in_cover[(size_t)v] = 1;
  leads.assign((size_t)P, 0);
  for (const auto& e : edges) leads[(size_t)(in_cover[(size_t)e.first] ? e.first : e.second)] = 1;
}

// The plan into `steps`, whose capacity is reused: a caller that plans
// again and again keeps one buffer and its pages.  `cover`, when given, is
// plan_cover's in_cover for the same integrals (checked there).
inline void make_plan_into(const int32_t* quads, int64_t n, int n_states, bool ordered,
                           std::vector<PlanStep>& steps, const std::vector<char>* cover = nullptr) {
  if (!cover) check_quads(quads, n, n_states);
  steps.resize((size_t)n);
  const int64_t P = (int64_t)n_states * (n_states + 1) / 2;
  parallel_ranges((size_t)n, [&](size_t k0, size_t k1, size_t) {
    for (size_t k = k0; k < k1; ++k) {
      const int32_t* q = quads + 4 * k;
      steps[k] = PlanStep{(int)k, plan_product(q[0], q[3]), plan_product(q[1], q[2])};
    }
  });
  std::vector<char> own;
  if (!cover) {
    const auto edges = product_edges(steps.size(), P, [&](size_t k) {
      return (uint64_t)steps[k].left.product * P + steps[k].right.product;
    });
    own.assign((size_t)P, 0);
    for (int v : vertex_cover((int)P, edges)) own[(size_t)v] = 1;
    cover = &own;
  }
  const std::vector<char>& in_cover = *cover;
  parallel_ranges(steps.size(), [&](size_t k0, size_t k1, size_t) {
    for (size_t k = k0; k < k1; ++k)
      if (!in_cover[(size_t)steps[k].left.product]) std::swap(steps[k].left, steps[k].right);
  });
  if (!ordered) return;
  // order (Planner.hpp:94-101): left product, right product, then the steps
  // whose two products have equal conj flags before the others
  auto rc = [](const PlanStep& s) { return (s.left.conj != s.right.conj) ? 1 : 0; };
  const int64_t n_keys = P * P * 2;
  if (n_keys <= std::max<int64_t>((int64_t)1 << 20, std::min<int64_t>(4 * n, (int64_t)1 << 25))) {
    // one stable counting pass over (left, right, conj flag): the steps are
    // in id order, so equal keys keep it
    std::vector<uint32_t> key(steps.size());
    std::vector<uint32_t> start((size_t)n_keys + 1, 0);
    for (size_t k = 0; k < steps.size(); ++k) {
      const PlanStep& s = steps[k];
      key[k] = (uint32_t)(((int64_t)s.left.product * P + s.right.product) * 2 + rc(s));
      ++start[key[k] + 1];
    }
    for (int64_t i = 0; i < n_keys; ++i) start[(size_t)i + 1] += start[(size_t)i];
    std::vector<PlanStep> sorted(steps.size());
    for (size_t k = 0; k < steps.size(); ++k) sorted[start[key[k]]++] = steps[k];
    std::copy(sorted.begin(), sorted.end(), steps.begin());
    return;
  }
  if (n_keys < ((int64_t)1 << 32)) {
    std::vector<uint64_t> key;
    key.reserve(steps.size());
    for (const PlanStep& s : steps)
      key.push_back((((uint64_t)s.left.product * P + s.right.product) * 2 + rc(s)) << 31 | (uint64_t)s.id);
    radix_sort_u64(key, 64);
    std::vector<PlanStep> sorted(steps.size());
    for (size_t i = 0; i < key.size(); ++i) sorted[i] = steps[(size_t)(key[i] & 0x7FFFFFFF)];
    std::copy(sorted.begin(), sorted.end(), steps.begin());
    return;
  }
  std::sort(steps.begin(), steps.end(), [&rc](const PlanStep& a, const PlanStep& b) {
    if (a.left.product != b.left.product) return a.left.product < b.left.product;
    if (a.right.product != b.right.product) return a.right.product < b.right.product;
    return rc(a) < rc(b);
  });
}

inline std::vector<PlanStep> make_plan(const int32_t* quads, int64_t n, int n_states, bool ordered = true) {
  std::vector<PlanStep> steps;
  make_plan_into(quads, n, n_states, ordered, steps);
  return steps;
}

}  // namespace host
}  // namespace clat
