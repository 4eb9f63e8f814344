// Host-side greedy constraint coloring for the Gauss-Seidel sweep
// (src/xpbd3d/coloring.py:26-89), in C++ so a flush of a 100k-row cloth
// colors in milliseconds instead of the reference's pure-Python loop.
//
// The output must equal the reference bit for bit, so the visiting order is
// the reference's: a stable sort of row ids by -(deg[a] + deg[b] (b >= 0)),
// then each row takes the smallest color free on both endpoints.  Per-body
// color sets are kept as 64-bit masks with an overflow list for bodies that
// reach more than 64 colors (dense contact piles).
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "xpbd.h"

namespace {

struct ColorSet {
  std::vector<uint64_t> words;  // bit c set <=> color c used at this body
  bool has(int c) const {
    size_t w = (size_t)c >> 6;
    return w < words.size() && ((words[w] >> (c & 63)) & 1u);
  }
  void add(int c) {
    size_t w = (size_t)c >> 6;
    if (w >= words.size()) words.resize(w + 1, 0);
    words[w] |= (uint64_t)1 << (c & 63);
  }
};

}  // namespace

extern "C" {

int xpbd_color_constraints(int n_bodies, const int* body_a, const int* body_b, int n_rows,
                           int* color_out) {
  if (n_rows <= 0) return 0;
  if (!body_a || !body_b || !color_out || n_bodies < 0) return -1;
  for (int j = 0; j < n_rows; ++j) {
    if (body_a[j] < 0 || body_a[j] >= n_bodies || body_b[j] >= n_bodies) return -1;
  }
  std::vector<int64_t> deg((size_t)n_bodies, 0);
  for (int j = 0; j < n_rows; ++j) {
    deg[(size_t)body_a[j]] += 1;
    if (body_b[j] >= 0) deg[(size_t)body_b[j]] += 1;
  }
  std::vector<int64_t> key((size_t)n_rows);
  for (int j = 0; j < n_rows; ++j)
    key[(size_t)j] = -(deg[(size_t)body_a[j]] + (body_b[j] >= 0 ? deg[(size_t)body_b[j]] : 0));
  std::vector<int> order((size_t)n_rows);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int p, int q) { return key[(size_t)p] < key[(size_t)q]; });

  std::vector<ColorSet> used((size_t)n_bodies);
  int num_colors = 0;
  for (int j : order) {
    const ColorSet& ua = used[(size_t)body_a[j]];
    const int b = body_b[j];
    int c = 0;
    if (b >= 0) {
      const ColorSet& ub = used[(size_t)b];
      // first zero bit of (ua | ub), word by word
      size_t nw = std::max(ua.words.size(), ub.words.size());
      size_t w = 0;
      for (; w < nw; ++w) {
        uint64_t m = (w < ua.words.size() ? ua.words[w] : 0) | (w < ub.words.size() ? ub.words[w] : 0);
        if (~m) {
          c = (int)(w * 64 + __builtin_ctzll(~m));
          break;
        }
      }
      if (w == nw) c = (int)(nw * 64);
    } else {
      size_t nw = ua.words.size(), w = 0;
      for (; w < nw; ++w) {
        uint64_t m = ua.words[w];
        if (~m) {
          c = (int)(w * 64 + __builtin_ctzll(~m));
          break;
        }
      }
      if (w == nw) c = (int)(nw * 64);
    }
    color_out[j] = c;
    used[(size_t)body_a[j]].add(c);
    if (b >= 0) used[(size_t)b].add(c);
    num_colors = std::max(num_colors, c + 1);
  }
  return num_colors;
}

// coloring.py:74-89 — order = stable argsort of color, starts = prefix sums.
int xpbd_color_ranges(const int* color, int n_rows, int num_colors, int* order_out,
                      int* starts_out) {
  if (n_rows <= 0) {
    if (starts_out) starts_out[0] = 0;
    return 0;
  }
  if (!color || !order_out || !starts_out || num_colors < 1) return -1;
  std::vector<int> count((size_t)num_colors + 1, 0);
  for (int j = 0; j < n_rows; ++j) {
    if (color[j] < 0 || color[j] >= num_colors) return -1;
    count[(size_t)color[j] + 1] += 1;
  }
  for (int c = 0; c < num_colors; ++c) count[(size_t)c + 1] += count[(size_t)c];
  for (int c = 0; c <= num_colors; ++c) starts_out[c] = count[(size_t)c];
  for (int j = 0; j < n_rows; ++j) order_out[count[(size_t)color[j]]++] = j;
  return 0;
}

}  // extern "C"
