// Seeded synthetic scene generators for the benchmark workloads and tests
// (host C++, its own small library: it produces inputs, it renders nothing).
//
// PRNG: SplitMix64, uniform doubles as (u >> 11) * 2^-53 -- portable across
// standard libraries, unlike std::uniform_*_distribution.
#include <algorithm>
#include <cmath>
#include <cstdint>
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
};

template <class F>
void parallel_slices(int nz, int threads, F &&fn) {
  if (threads <= 0) threads = int(std::max(1u, std::thread::hardware_concurrency()));
  threads = std::min(threads, nz);
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&, t]() {
      for (int z = t; z < nz; z += threads) fn(z);
    });
  for (auto &th : pool) th.join();
}

// Improved gradient noise (Perlin 2002) over a seeded permutation.
struct Noise {
  int p[512];
  explicit Noise(uint64_t seed) {
    SplitMix64 r(seed);
    int perm[256];
    for (int i = 0; i < 256; ++i) perm[i] = i;
    for (int i = 255; i > 0; --i) {  // Fisher-Yates
      int j = int(r.next() % uint64_t(i + 1));
      std::swap(perm[i], perm[j]);
    }
    for (int i = 0; i < 512; ++i) p[i] = perm[i & 255];
  }
  static double fade(double t) { return t * t * t * (t * (t * 6 - 15) + 10); }
  static double lerp(double t, double a, double b) { return a + t * (b - a); }
  static double grad(int h, double x, double y, double z) {
    h &= 15;
    double u = h < 8 ? x : y;
    double v = h < 4 ? y : (h == 12 || h == 14 ? x : z);
    return ((h & 1) ? -u : u) + ((h & 2) ? -v : v);
  }
  double operator()(double x, double y, double z) const {
    double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
    int X = int(int64_t(fx) & 255), Y = int(int64_t(fy) & 255), Z = int(int64_t(fz) & 255);
    x -= fx;
    y -= fy;
    z -= fz;
    double u = fade(x), v = fade(y), w = fade(z);
    int A = p[X] + Y, AA = p[A] + Z, AB = p[A + 1] + Z;
    int B = p[X + 1] + Y, BA = p[B] + Z, BB = p[B + 1] + Z;
    return lerp(w,
                lerp(v, lerp(u, grad(p[AA], x, y, z), grad(p[BA], x - 1, y, z)),
                     lerp(u, grad(p[AB], x, y - 1, z), grad(p[BB], x - 1, y - 1, z))),
                lerp(v, lerp(u, grad(p[AA + 1], x, y, z - 1), grad(p[BA + 1], x - 1, y, z - 1)),
                     lerp(u, grad(p[AB + 1], x, y - 1, z - 1),
                          grad(p[BB + 1], x - 1, y - 1, z - 1))));
  }
};

}  // namespace

extern "C" {

__attribute__((visibility("default"))) void vrtgen_splitmix_uniform(uint64_t seed, int n,
                                                                     double *out) {
  SplitMix64 r(seed);
  for (int i = 0; i < n; ++i) out[i] = r.uniform();
}

// Sum of `count` isotropic Gaussian blobs sampled at the voxel positions of a
// field spanning [-0.5, 0.5]^3 (voxel i at -0.5 + i/(N-1)), clamped to [0,1].
// Centres ~ U[0,1]^3 (unit cube, shifted onto the bounds), sigma ~
// U[0.03,0.12], amplitude ~ U[0.2,1], drawn in that order per blob.
__attribute__((visibility("default"))) void vrtgen_blobs(const int dims[3], int count,
                                                          uint64_t seed, float *out,
                                                          int threads) {
  SplitMix64 r(seed);
  std::vector<double> cx(count), cy(count), cz(count), s2(count), amp(count);
  for (int i = 0; i < count; ++i) {
    cx[i] = r.uniform() - 0.5;
    cy[i] = r.uniform() - 0.5;
    cz[i] = r.uniform() - 0.5;
    double sigma = 0.03 + 0.09 * r.uniform();
    s2[i] = 1.0 / (2.0 * sigma * sigma);
    amp[i] = 0.2 + 0.8 * r.uniform();
  }
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  parallel_slices(nz, threads, [&](int z) {
    double pz = -0.5 + (nz > 1 ? double(z) / (nz - 1) : 0.5);
    for (int y = 0; y < ny; ++y) {
      double py = -0.5 + (ny > 1 ? double(y) / (ny - 1) : 0.5);
      for (int x = 0; x < nx; ++x) {
        double px = -0.5 + (nx > 1 ? double(x) / (nx - 1) : 0.5);
        double v = 0;
        for (int i = 0; i < count; ++i) {
          double dx = px - cx[i], dy = py - cy[i], dz = pz - cz[i];
          v += amp[i] * std::exp(-(dx * dx + dy * dy + dz * dz) * s2[i]);
        }
        out[(size_t(z) * ny + y) * nx + x] = float(std::min(std::max(v, 0.0), 1.0));
      }
    }
  });
}

// fBm of improved gradient noise, evaluated at unit-cube voxel positions
// q = i/(N-1) scaled by `base_freq`, `octaves` octaves with lacunarity and
// gain, remapped to [0,1] by its own min/max.
__attribute__((visibility("default"))) void vrtgen_fbm(const int dims[3], int octaves,
                                                        double base_freq, double lacunarity,
                                                        double gain, uint64_t seed, float *out,
                                                        int threads) {
  Noise noise(seed);
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  std::vector<double> zmin(nz, 1e300), zmax(nz, -1e300);
  parallel_slices(nz, threads, [&](int z) {
    double qz = nz > 1 ? double(z) / (nz - 1) : 0.5;
    for (int y = 0; y < ny; ++y) {
      double qy = ny > 1 ? double(y) / (ny - 1) : 0.5;
      for (int x = 0; x < nx; ++x) {
        double qx = nx > 1 ? double(x) / (nx - 1) : 0.5;
        double f = base_freq, a = 1.0, v = 0.0;
        for (int o = 0; o < octaves; ++o) {
          v += a * noise(qx * f, qy * f, qz * f);
          f *= lacunarity;
          a *= gain;
        }
        out[(size_t(z) * ny + y) * nx + x] = float(v);
        zmin[z] = std::min(zmin[z], v);
        zmax[z] = std::max(zmax[z], v);
      }
    }
  });
  double lo = *std::min_element(zmin.begin(), zmin.end());
  double hi = *std::max_element(zmax.begin(), zmax.end());
  double inv = hi > lo ? 1.0 / (hi - lo) : 0.0;
  parallel_slices(nz, threads, [&](int z) {
    float *row = out + size_t(z) * ny * nx;
    for (size_t i = 0; i < size_t(ny) * nx; ++i) {
      double v = (double(row[i]) - lo) * inv;
      row[i] = float(std::min(std::max(v, 0.0), 1.0));
    }
  });
}

}  // extern "C"
