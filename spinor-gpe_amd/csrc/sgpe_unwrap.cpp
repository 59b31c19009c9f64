// 2D phase unwrapping, reliability-sorted, non-continuous path
// (M. A. Herraez, D. R. Burton, M. J. Lalor, M. A. Gdeisat, "Fast
// two-dimensional phase-unwrapping algorithm based on sorting by reliability
// following a noncontinuous path", Appl. Opt. 41, 7437-7444 (2002)).
//
// This is the published algorithm behind skimage.restoration.unwrap_phase,
// which the reference calls for its end-of-run energy
// (spinor_gpe/pspinor/tensor_tools.py:528-531). Steps:
//  1. reliability of every interior pixel from the wrapped second
//     differences along x, y and both diagonals, D = sqrt(H^2+V^2+D1^2+D2^2)
//     (border pixels get a large constant D: least reliable);
//  2. every horizontal and vertical edge gets D1 + D2 of its pixels and the
//     2 pi jump between them;
//  3. edges are visited from most to least reliable; the two pixels' groups
//     are merged, the smaller group (pixel 1's group on a tie) being shifted
//     by whole turns so that the edge is continuous;
//  4. every pixel gets its accumulated turns added.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/sgpe.h"

namespace {

const double kTwoPi = 6.283185307179586476925;

inline double wrap(double d) {  // into (-pi, pi]
  if (d > M_PI) return d - kTwoPi;
  if (d < -M_PI) return d + kTwoPi;
  return d;
}

inline int jump(double p1, double p2) {  // turns to add to p2 relative to p1
  const double d = p1 - p2;
  if (d > M_PI) return 1;
  if (d < -M_PI) return -1;
  return 0;
}

struct Edge {
  double key;
  int32_t p1, p2;
  int32_t inc;
};

}  // namespace

static int unwrap_2d(const double* ph, double* out, int ny, int nx) {
  const int n = ny * nx;
  std::vector<double> rel(n, 1.0e7);
  for (int i = 1; i + 1 < ny; ++i)
    for (int j = 1; j + 1 < nx; ++j) {
      const double c = ph[i * nx + j];
      const double h = wrap(ph[i * nx + j - 1] - c) - wrap(c - ph[i * nx + j + 1]);
      const double v = wrap(ph[(i - 1) * nx + j] - c) - wrap(c - ph[(i + 1) * nx + j]);
      const double d1 = wrap(ph[(i - 1) * nx + j - 1] - c) - wrap(c - ph[(i + 1) * nx + j + 1]);
      const double d2 = wrap(ph[(i - 1) * nx + j + 1] - c) - wrap(c - ph[(i + 1) * nx + j - 1]);
      rel[i * nx + j] = std::sqrt(h * h + v * v + d1 * d1 + d2 * d2);
    }
  std::vector<Edge> edges;
  edges.reserve((size_t)2 * n);
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < nx; ++j) {
      const int p = i * nx + j;
      if (j + 1 < nx) edges.push_back({rel[p] + rel[p + 1], p, p + 1, jump(ph[p], ph[p + 1])});
      if (i + 1 < ny) edges.push_back({rel[p] + rel[p + nx], p, p + nx, jump(ph[p], ph[p + nx])});
    }
  std::stable_sort(edges.begin(), edges.end(), [](const Edge& a, const Edge& b) { return a.key < b.key; });

  // Groups as linked lists: head[p] = group id (its first pixel), next[p],
  // last[g], size[g], turns[p].
  std::vector<int32_t> head(n), next(n, -1), last(n), size(n, 1);
  std::vector<int32_t> turns(n, 0);
  for (int p = 0; p < n; ++p) {
    head[p] = p;
    last[p] = p;
  }
  for (const Edge& e : edges) {
    const int g1 = head[e.p1], g2 = head[e.p2];
    if (g1 == g2) continue;
    int src, dst, shift;
    if (size[g1] > size[g2]) {  // move pixel 2's group next to pixel 1's
      dst = g1;
      src = g2;
      shift = turns[e.p1] + e.inc - turns[e.p2];
    } else {  // move pixel 1's group
      dst = g2;
      src = g1;
      shift = turns[e.p2] - e.inc - turns[e.p1];
    }
    for (int q = src; q != -1; q = next[q]) {
      head[q] = dst;
      turns[q] += shift;
    }
    next[last[dst]] = src;
    last[dst] = last[src];
    size[dst] += size[src];
  }
  for (int p = 0; p < n; ++p) out[p] = ph[p] + kTwoPi * turns[p];
  return 0;
}

extern "C" int sgpe_unwrap_phase_2d(const double* ph, double* out, int ny, int nx) {
  if (!ph || !out || ny < 1 || nx < 1 || (int64_t)ny * nx > INT32_MAX) return SGPE_E_ARG;
  try {
    return unwrap_2d(ph, out, ny, nx);
  } catch (...) {  // std::bad_alloc of the work arrays: no exception crosses the C ABI
    return SGPE_E_NOMEM;
  }
}
