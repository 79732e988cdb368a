// Seeded synthetic volumes (declarations and contract: include/cvr_synth.h).
#include "../../../include/cvr_synth.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct SplitMix64 {
  uint64_t s;
  explicit SplitMix64(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uniform() { return double(next() >> 11) * (1.0 / 9007199254740992.0); }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

template <class F> void parallel_for(int n, F&& f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4 || nt == 1) {
    for (int i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&, t] {
      for (int i = int(t); i < n; i += int(nt)) f(i);
    });
  for (auto& x : th) x.join();
}

// Value noise: lattice values drawn per octave in (z, y, x) order.
struct ValueNoise {
  int octaves;
  std::vector<int> res;
  std::vector<std::vector<float>> lat;
  ValueNoise(int oct, SplitMix64& rng) : octaves(oct) {
    for (int o = 0; o < oct; ++o) {
      int L = 8 << o;
      res.push_back(L);
      std::vector<float> v(size_t(L + 1) * (L + 1) * (L + 1));
      for (auto& x : v) x = float(rng.uniform());
      lat.push_back(std::move(v));
    }
  }
  static double fade(double t) { return t * t * (3.0 - 2.0 * t); }
  double eval(double u, double v, double w) const {
    double sum = 0, norm = 0, amp = 1;
    for (int o = 0; o < octaves; ++o) {
      int L = res[o];
      double p[3] = {u * L, v * L, w * L};
      int i[3];
      double t[3];
      for (int k = 0; k < 3; ++k) {
        double fl = std::floor(p[k]);
        i[k] = std::min(std::max(int(fl), 0), L - 1);
        t[k] = fade(std::min(std::max(p[k] - i[k], 0.0), 1.0));
      }
      const std::vector<float>& g = lat[o];
      auto at = [&](int x, int y, int z) { return double(g[size_t(x) + size_t(L + 1) * (y + size_t(L + 1) * z)]); };
      double c00 = at(i[0], i[1], i[2]) * (1 - t[0]) + at(i[0] + 1, i[1], i[2]) * t[0];
      double c10 = at(i[0], i[1] + 1, i[2]) * (1 - t[0]) + at(i[0] + 1, i[1] + 1, i[2]) * t[0];
      double c01 = at(i[0], i[1], i[2] + 1) * (1 - t[0]) + at(i[0] + 1, i[1], i[2] + 1) * t[0];
      double c11 = at(i[0], i[1] + 1, i[2] + 1) * (1 - t[0]) + at(i[0] + 1, i[1] + 1, i[2] + 1) * t[0];
      double c0 = c00 * (1 - t[1]) + c10 * t[1], c1 = c01 * (1 - t[1]) + c11 * t[1];
      sum += amp * (c0 * (1 - t[2]) + c1 * t[2]);
      norm += amp;
      amp *= 0.5;
    }
    return sum / norm;
  }
};

void store(const std::vector<float>& f, int dtype, double scale_to, void* out) {
  const size_t n = f.size();
  if (dtype == 0) {
    uint8_t* o = static_cast<uint8_t*>(out);
    for (size_t i = 0; i < n; ++i) o[i] = uint8_t(std::lround(std::min(std::max(double(f[i]), 0.0), 1.0) * scale_to));
  } else {
    uint16_t* o = static_cast<uint16_t*>(out);
    for (size_t i = 0; i < n; ++i) o[i] = uint16_t(std::lround(std::min(std::max(double(f[i]), 0.0), 1.0) * scale_to));
  }
}

}  // namespace

extern "C" int cvr_synth_blobs(int nx, int ny, int nz, int dtype, int n_blobs, double noise_weight,
                               int noise_octaves, uint64_t seed, void* out) {
  if (nx < 1 || ny < 1 || nz < 1 || !out || n_blobs < 0 || (dtype != 0 && dtype != 1)) return -1;
  SplitMix64 rng(seed);
  struct Blob { double c[3], s, a; };
  std::vector<Blob> blobs(n_blobs);
  for (auto& b : blobs) {
    for (double& c : b.c) c = rng.uniform(0.2, 0.8);
    b.s = rng.uniform(0.05, 0.15);
    b.a = rng.uniform(0.3, 1.0);
  }
  const size_t sxy = size_t(nx) * ny;
  std::vector<float> f(sxy * nz, 0.f);
  const int n[3] = {nx, ny, nz};
  // separable Gaussians: g_axis[i] = exp(-(u_i - c)^2 / (2 s^2))
  std::vector<std::vector<double>> g(n_blobs * 3);
  for (int b = 0; b < n_blobs; ++b)
    for (int k = 0; k < 3; ++k) {
      g[b * 3 + k].resize(n[k]);
      for (int i = 0; i < n[k]; ++i) {
        double u = (i + 0.5) / n[k] - blobs[b].c[k];
        g[b * 3 + k][i] = std::exp(-(u * u) / (2.0 * blobs[b].s * blobs[b].s));
      }
    }
  ValueNoise* vn = nullptr;
  SplitMix64 nrng(seed ^ 0x5DEECE66Dull);
  if (noise_weight != 0 && noise_octaves > 0) vn = new ValueNoise(noise_octaves, nrng);
  parallel_for(nz, [&](int z) {
    for (int y = 0; y < ny; ++y) {
      float* row = f.data() + size_t(z) * sxy + size_t(y) * nx;
      for (int b = 0; b < n_blobs; ++b) {
        double zy = blobs[b].a * g[b * 3 + 2][z] * g[b * 3 + 1][y];
        if (zy < 1e-12) continue;
        const double* gx = g[b * 3].data();
        for (int x = 0; x < nx; ++x) row[x] += float(zy * gx[x]);
      }
      if (vn)
        for (int x = 0; x < nx; ++x)
          row[x] += float(noise_weight * vn->eval((x + 0.5) / nx, (y + 0.5) / ny, (z + 0.5) / nz));
    }
  });
  delete vn;
  float mx = 0.f;
  for (float v : f) mx = std::max(mx, v);
  if (mx > 0)
    for (float& v : f) v /= mx;
  store(f, dtype, dtype == 0 ? 255.0 : 65535.0, out);
  return 0;
}

extern "C" int cvr_synth_noise(int nx, int ny, int nz, int dtype, int octaves, uint64_t seed, int n_bands,
                               const double* band_values, void* out) {
  if (nx < 1 || ny < 1 || nz < 1 || !out || octaves < 1 || (dtype != 0 && dtype != 1)) return -1;
  if (n_bands > 0 && !band_values) return -1;
  SplitMix64 rng(seed);
  ValueNoise vn(octaves, rng);
  const size_t sxy = size_t(nx) * ny;
  std::vector<float> f(sxy * nz);
  parallel_for(nz, [&](int z) {
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        f[size_t(z) * sxy + size_t(y) * nx + x] = float(vn.eval((x + 0.5) / nx, (y + 0.5) / ny, (z + 0.5) / nz));
  });
  if (n_bands > 0) {
    std::vector<float> sorted(f);
    std::vector<float> cuts(n_bands - 1);
    for (int b = 1; b < n_bands; ++b) {
      size_t k = size_t(double(sorted.size() - 1) * b / n_bands);
      std::nth_element(sorted.begin(), sorted.begin() + k, sorted.end());
      cuts[b - 1] = sorted[k];
    }
    for (float& v : f) {
      int b = 0;
      while (b < n_bands - 1 && v > cuts[b]) ++b;
      v = float(band_values[b]);
    }
  } else {
    float lo = *std::min_element(f.begin(), f.end()), hi = *std::max_element(f.begin(), f.end());
    for (float& v : f) v = hi > lo ? (v - lo) / (hi - lo) : 0.f;
  }
  store(f, dtype, dtype == 0 ? 255.0 : 65535.0, out);
  return 0;
}
