// TEST INFRASTRUCTURE ONLY.  Thin extern "C" driver around the reference's
// own self-contained headers, compiled from where they lie in the reference
// tree (src/Round.hpp, src/Graph.hpp, src/Pattern.hpp + src/throwf.hpp).
// The result, oracle/_ref/libref.so, is the checker for the host-side
// integer logic; nothing of it ships in the product.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#include "throwf.hpp"
#include "Round.hpp"
#include "Graph.hpp"
#include "Pattern.hpp"

extern "C" {

int ref_round_up_2(int n, int c) { return Round::up<2>(n, c); }
int ref_round_up_235(int n, int c) { return Round::up<2, 3, 5>(n, c); }

// cover: capacity nv; returns the cover size
int ref_vertex_cover(int nv, const int32_t* edges, int64_t ne, int32_t* cover) {
  Graph g(nv);
  for (int64_t k = 0; k < ne; ++k) g.addEdge(edges[2 * k], edges[2 * k + 1]);
  auto c = g.computeVertexCover();
  int i = 0;
  for (int v : c) cover[i++] = v;
  return i;
}

// Enumerates matches of the pattern list in coulombo.cpp's loop order.
// Patterns must have at least four characters (shorter ones read past the
// end of the string inside the reference).  Returns the count, or -1 with
// the exception text in err.
int64_t ref_pattern_enumerate(const char* spec, int n_states, int n_holes, int32_t* quads, int64_t cap,
                              char* err, int errlen) {
  try {
    Pattern p(spec, n_holes);
    int64_t k = 0;
    for (int i1 = 1; i1 <= n_states; ++i1)
      for (int i2 = 1; i2 <= n_states; ++i2)
        for (int i3 = 1; i3 <= n_states; ++i3)
          for (int i4 = 1; i4 <= n_states; ++i4)
            if (p.match(i1, i2, i3, i4)) {
              if (k < cap) {
                quads[4 * k] = i1; quads[4 * k + 1] = i2; quads[4 * k + 2] = i3; quads[4 * k + 3] = i4;
              }
              ++k;
            }
    return k;
  } catch (const std::exception& e) {
    if (err && errlen > 0) {
      std::strncpy(err, e.what(), errlen - 1);
      err[errlen - 1] = 0;
    }
    return -1;
  }
}

}
