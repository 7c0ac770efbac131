// Seeded synthetic meshes and poses (see include/fclx_synth.h).
#include "../../../include/fclx_synth.h"

#include <array>
#include <cmath>
#include <map>
#include <tuple>
#include <vector>

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t z = seed;
    for (auto& w : s) {  // splitmix64
      z += 0x9e3779b97f4a7c15ULL;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
      x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
      w = x ^ (x >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  double normal() {
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

void rotation_from_quat(double w, double x, double y, double z, double* R) {
  const double n = std::sqrt(w * w + x * x + y * y + z * z);
  w /= n;
  x /= n;
  y /= n;
  z /= n;
  R[0] = 1 - 2 * (y * y + z * z);
  R[1] = 2 * (x * y - w * z);
  R[2] = 2 * (x * z + w * y);
  R[3] = 2 * (x * y + w * z);
  R[4] = 1 - 2 * (x * x + z * z);
  R[5] = 2 * (y * z - w * x);
  R[6] = 2 * (x * z - w * y);
  R[7] = 2 * (y * z + w * x);
  R[8] = 1 - 2 * (x * x + y * y);
}

const double kPhi = 1.6180339887498948482;
const double kIcoV[12][3] = {{-1, kPhi, 0}, {1, kPhi, 0},  {-1, -kPhi, 0}, {1, -kPhi, 0},
                             {0, -1, kPhi}, {0, 1, kPhi},  {0, -1, -kPhi}, {0, 1, -kPhi},
                             {kPhi, 0, -1}, {kPhi, 0, 1},  {-kPhi, 0, -1}, {-kPhi, 0, 1}};
const int kIcoF[20][3] = {{0, 11, 5}, {0, 5, 1},   {0, 1, 7},   {0, 7, 10}, {0, 10, 11},
                          {1, 5, 9},  {5, 11, 4},  {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                          {3, 9, 4},  {3, 4, 2},   {3, 2, 6},   {3, 6, 8},  {3, 8, 9},
                          {4, 9, 5},  {2, 4, 11},  {6, 2, 10},  {8, 6, 7},  {9, 8, 1}};

}  // namespace

extern "C" {

void fclx_synth_icosphere_size(int32_t k, int32_t* nv, int32_t* nt) {
  *nv = 10 * k * k + 2;
  *nt = 20 * k * k;
}

int fclx_synth_icosphere(int32_t k, double noise, uint64_t seed, double* vertices,
                         int32_t* tris) {
  if (k < 1) return -1;
  std::vector<std::array<double, 3>> V;
  V.reserve(10 * static_cast<size_t>(k) * k + 2);
  for (auto& c : kIcoV) V.push_back({c[0], c[1], c[2]});
  // edge points: key (lo, hi, step from lo)
  std::map<std::tuple<int, int, int>, int> edge_ids;
  auto lerp_add = [&](const double* a, const double* b, const double* c, int i, int j) {
    std::array<double, 3> p;
    for (int d = 0; d < 3; ++d)
      p[d] = a[d] + (b[d] - a[d]) * i / static_cast<double>(k) +
             (c[d] - a[d]) * j / static_cast<double>(k);
    V.push_back(p);
    return static_cast<int>(V.size()) - 1;
  };
  auto edge_point = [&](int u, int v, int t) {  // t steps from u toward v
    if (t == 0) return u;
    if (t == k) return v;
    const int lo = u < v ? u : v, hi = u < v ? v : u;
    const int tt = u < v ? t : k - t;
    auto key = std::make_tuple(lo, hi, tt);
    auto it = edge_ids.find(key);
    if (it != edge_ids.end()) return it->second;
    const double* a = kIcoV[lo];
    const double* b = kIcoV[hi];
    int id = lerp_add(a, b, a, tt, 0);
    edge_ids[key] = id;
    return id;
  };
  int nt = 0;
  for (const auto& f : kIcoF) {
    const int a = f[0], b = f[1], c = f[2];
    // index grid for this face: (i, j) with i + j <= k
    std::vector<int> grid((k + 1) * (k + 1), -1);
    auto at = [&](int i, int j) -> int& { return grid[i * (k + 1) + j]; };
    for (int i = 0; i <= k; ++i) {
      for (int j = 0; i + j <= k; ++j) {
        int id;
        if (j == 0) id = edge_point(a, b, i);
        else if (i == 0) id = edge_point(a, c, j);
        else if (i + j == k) id = edge_point(b, c, j);
        else id = lerp_add(kIcoV[a], kIcoV[b], kIcoV[c], i, j);
        at(i, j) = id;
      }
    }
    for (int i = 0; i < k; ++i) {
      for (int j = 0; i + j < k; ++j) {
        tris[3 * nt] = at(i, j);
        tris[3 * nt + 1] = at(i + 1, j);
        tris[3 * nt + 2] = at(i, j + 1);
        ++nt;
        if (i + j < k - 1) {
          tris[3 * nt] = at(i + 1, j);
          tris[3 * nt + 1] = at(i + 1, j + 1);
          tris[3 * nt + 2] = at(i, j + 1);
          ++nt;
        }
      }
    }
  }
  Rng rng(seed);
  for (size_t i = 0; i < V.size(); ++i) {
    const auto& p = V[i];
    const double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double r = 1.0 + noise * (2.0 * rng.uniform() - 1.0);
    for (int d = 0; d < 3; ++d) vertices[3 * i + d] = p[d] / n * r;
  }
  return 0;
}

void fclx_synth_heightfield_size(int32_t g, int32_t* nv, int32_t* nt) {
  *nv = g * g;
  *nt = 2 * (g - 1) * (g - 1);
}

int fclx_synth_heightfield(int32_t g, double height_frac, uint64_t seed, double* vertices,
                           int32_t* tris) {
  if (g < 2) return -1;
  Rng rng(seed);
  const double h = height_frac * 2.0;
  for (int i = 0; i < g; ++i)
    for (int j = 0; j < g; ++j) {
      double* p = vertices + 3 * (static_cast<size_t>(i) * g + j);
      p[0] = -1.0 + 2.0 * i / (g - 1);
      p[1] = -1.0 + 2.0 * j / (g - 1);
      p[2] = h * (2.0 * rng.uniform() - 1.0);
    }
  int nt = 0;
  for (int i = 0; i + 1 < g; ++i)
    for (int j = 0; j + 1 < g; ++j) {
      const int v00 = i * g + j, v10 = (i + 1) * g + j, v11 = (i + 1) * g + j + 1,
                v01 = i * g + j + 1;
      tris[3 * nt] = v00;
      tris[3 * nt + 1] = v10;
      tris[3 * nt + 2] = v11;
      ++nt;
      tris[3 * nt] = v00;
      tris[3 * nt + 1] = v11;
      tris[3 * nt + 2] = v01;
      ++nt;
    }
  return 0;
}

void fclx_synth_scale_to_radius(double* v, int32_t nv, double r) {
  double mx = 0;
  for (int i = 0; i < nv; ++i) {
    const double* p = v + 3 * static_cast<size_t>(i);
    mx = std::fmax(mx, std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]));
  }
  if (mx <= 0) return;
  const double s = r / mx;
  for (int64_t i = 0; i < 3LL * nv; ++i) v[i] *= s;
}

int fclx_synth_poses(int32_t n, int32_t kind, double p0, double p1, uint64_t seed,
                     double* tf1, double* tf2) {
  if (n < 0 || (kind != 0 && kind != 1)) return -1;
  Rng rng(seed);
  for (int q = 0; q < n; ++q) {
    double* a = tf1 + 12 * static_cast<size_t>(q);
    double* b = tf2 + 12 * static_cast<size_t>(q);
    for (int i = 0; i < 12; ++i) a[i] = 0;
    a[0] = a[4] = a[8] = 1;
    const double w = rng.normal(), x = rng.normal(), y = rng.normal(), z = rng.normal();
    rotation_from_quat(w, x, y, z, b);
    if (kind == 0) {
      for (int d = 0; d < 3; ++d) b[9 + d] = p0 * (2.0 * rng.uniform() - 1.0);
    } else {
      double u[3] = {rng.normal(), rng.normal(), rng.normal()};
      const double un = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
      const double s = p0 + (p1 - p0) * rng.uniform();
      for (int d = 0; d < 3; ++d) b[9 + d] = s * u[d] / un;
    }
  }
  return 0;
}

void fclx_synth_uniform(int64_t n, uint64_t seed, double* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = rng.uniform();
}

}  // extern "C"
