// median_fill.cpp -- the depth hole fills of src/recursiveMedianFiltering.cpp,
// on the host (they are sequential by construction: each filled pixel changes
// the neighbour counts that order the next choice).
//
// recursiveMedianFilter keeps the reference's container and insert/erase
// sequence (std::unordered_set of row-major indices, :44-98) so that, with the
// same standard library, pixels are visited in the same order.  The fill
// values pass through a float buffer as in the reference (:64,84,95).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "../../include/rgbdflow/OpticalFlow.h"

namespace {

// boost::hash_combine over (first, second) as the reference's std::hash<pair>
// specialisation uses it (include/rgbd_util/mathUtils.h:43).
struct PairHash {
  size_t operator()(const std::pair<int32_t, int32_t>& p) const {
    size_t seed = 0;
    seed ^= std::hash<int32_t>()(p.first) + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    seed ^= std::hash<int32_t>()(p.second) + 0x9e3779b9 + (seed << 6) + (seed >> 2);
    return seed;
  }
};

}  // namespace

// src/recursiveMedianFiltering.cpp:34-120
void recursiveMedianFilter(DImage& depth, const int32_t hw) {
  const int W = depth.width(), H = depth.height();
  double* data = depth.data();
  const int win = (2 * hw + 1) * (2 * hw + 1);
  std::vector<std::unordered_set<uint32_t>> byCount(win);
  std::vector<int32_t> countOf((size_t)W * H, -1);
  uint32_t left = 0;
  for (int32_t i = 0, l = 0; i < H; i++)
    for (int32_t j = 0; j < W; j++, l++)
      if (data[l] <= 0) {
        uint32_t c = 0;
        for (int32_t ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
          for (int32_t jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++)
            if (data[ii * W + jj] > 0) c++;
        byCount[c].insert(l);
        countOf[l] = c;
        left++;
      }
  std::vector<float> vals(win);
  while (left > 0) {
    int32_t mx = (int32_t)byCount.size() - 1;
    while (mx >= 0 && byCount[mx].empty()) mx--;
    if (mx <= 0) throw std::runtime_error("recursiveMedianFilter: an invalid region has no valid neighbour");
    const uint32_t l = *byCount[mx].begin();
    const int32_t i = l / W, j = l % W;
    byCount[mx].erase(l);
    countOf[l] = -1;
    uint32_t idx = 0;
    for (int32_t ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
      for (int32_t jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++) {
        const int32_t ll = ii * W + jj;
        if (data[ll] > 0) {
          vals[idx++] = data[ll];
        } else if (countOf[ll] >= 0) {
          byCount[countOf[ll]].erase(ll);
          countOf[ll]++;
          byCount[countOf[ll]].insert(ll);
        }
      }
    std::sort(vals.begin(), vals.begin() + idx);
    data[l] = vals[(int)std::floor(idx / 2)];
    left--;
  }
}

// src/recursiveMedianFiltering.cpp:127-222: fill only invalid 4-connected
// regions of at most maxRegionSizeToFill pixels, each pixel by the mean of its
// valid 3x3 neighbours weighted by exp(-|colour difference|^2 / 0.6).
void recursiveMedianFilterSmallInvalidRegionsUsingColor(DImage& depth, const DImage& col,
                                                        const uint32_t maxRegionSizeToFill) {
  typedef std::pair<int32_t, int32_t> ptT;  // (x, y)
  const int W = depth.width(), H = depth.height();
  const int32_t hw = 1;
  // connected components of invalid pixels (union-find over row-major indices)
  std::vector<int32_t> parent((size_t)W * H, -1);
  auto find = [&](int32_t a) {
    while (parent[a] != a) {
      parent[a] = parent[parent[a]];
      a = parent[a];
    }
    return a;
  };
  for (int32_t i = 0; i < H; i++)
    for (int32_t j = 0; j < W; j++)
      if (depth(i, j) <= 0) parent[i * W + j] = i * W + j;
  for (int32_t i = 0; i < H; i++)
    for (int32_t j = 0; j < W; j++)
      if (depth(i, j) <= 0) {
        const int32_t a = i * W + j;
        if (i > 0 && depth(i - 1, j) <= 0) parent[find(a)] = find(a - W);
        if (j > 0 && depth(i, j - 1) <= 0) parent[find(a)] = find(a - 1);
      }
  std::unordered_map<int32_t, std::vector<ptT>> byComp;
  std::vector<int32_t> roots;
  for (int32_t i = 0; i < H; i++)
    for (int32_t j = 0; j < W; j++)
      if (depth(i, j) <= 0) {
        const int32_t r = find(i * W + j);
        auto it = byComp.find(r);
        if (it == byComp.end()) {
          roots.push_back(r);
          it = byComp.emplace(r, std::vector<ptT>()).first;
        }
        it->second.push_back(std::make_pair(j, i));
      }
  for (int32_t r : roots) {
    const std::vector<ptT>& comp = byComp[r];
    if (comp.size() > maxRegionSizeToFill) continue;
    std::vector<std::unordered_set<ptT, PairHash>> byCount((2 * hw + 1) * (2 * hw + 1));
    std::unordered_map<ptT, int32_t, PairHash> countOf;
    uint32_t left = 0;
    for (const ptT& p : comp) {
      const int32_t i = p.second, j = p.first;
      uint32_t c = 0;
      for (int32_t ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
        for (int32_t jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++)
          if (depth(ii, jj) > 0) c++;
      byCount[c].insert(p);
      countOf[p] = c;
      left++;
    }
    while (left > 0) {
      int32_t mx = (int32_t)byCount.size() - 1;
      while (mx >= 0 && byCount[mx].empty()) mx--;
      if (mx <= 0) break;  // a region with no valid neighbour (the reference asserts here)
      const ptT p = *byCount[mx].begin();
      const int32_t i = p.second, j = p.first;
      byCount[mx].erase(p);
      countOf[p] = -1;
      float zSum = 0, weightSum = 0;
      for (int32_t ii = std::max(0, i - hw); ii <= std::min(H - 1, i + hw); ii++)
        for (int32_t jj = std::max(0, j - hw); jj <= std::min(W - 1, j + hw); jj++) {
          const ptT p2 = std::make_pair(jj, ii);
          if (depth(ii, jj) > 0) {
            const double d0 = col(i, j, 0) - col(ii, jj, 0), d1 = col(i, j, 1) - col(ii, jj, 1),
                         d2 = col(i, j, 2) - col(ii, jj, 2);
            const float w = std::exp(-(d0 * d0 + d1 * d1 + d2 * d2) / .6);
            zSum += depth(ii, jj) * w;
            weightSum += w;
          } else {
            auto it = countOf.find(p2);
            if (it != countOf.end() && it->second >= 0) {
              byCount[it->second].erase(p2);
              it->second++;
              byCount[it->second].insert(p2);
            }
          }
        }
      depth(i, j) = zSum / weightSum;
      left--;
    }
  }
}
