// workloads.cpp -- seeded synthetic meshes for tests and benchmarks.
// Host-only; see include/lbvh_workloads.h for the definitions.
#include "../../include/lbvh_workloads.h"
#include "synth_math.h"

#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

namespace {

constexpr float kPad = 3.1415f;  // parse-obj.js:56-58

using lbvh_synth::splitmix64;

// Uniform double in [0,1) for draw k of element i under `seed`.
inline double u01(uint64_t seed, uint64_t i, uint64_t k) {
  return lbvh_synth::u01(splitmix64(seed), i, k);
}

// Standard normal from draws k and k+1 (Box-Muller), bit-identical to the
// device's jitter (k_synth.hip).
inline double gauss(uint64_t seed, uint64_t i, uint64_t k) {
  return lbvh_synth::gauss(splitmix64(seed), i, k);
}

template <class F>
void parallel_for(uint64_t n, F f) {
  unsigned hw = std::thread::hardware_concurrency();
  unsigned nt = std::max(1u, std::min(hw ? hw : 1u, 16u));
  if (n < 65536) nt = 1;
  std::vector<std::thread> ts;
  const uint64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const uint64_t b = t * chunk, e = std::min(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back([=] {
      for (uint64_t i = b; i < e; ++i) f(i);
    });
  }
  for (auto &th : ts) th.join();
}

inline void put_tri(float *t, const double v[9]) {
  for (int k = 0; k < 3; ++k) {
    t[4 * k + 0] = (float)v[3 * k + 0];
    t[4 * k + 1] = (float)v[3 * k + 1];
    t[4 * k + 2] = (float)v[3 * k + 2];
    t[4 * k + 3] = kPad;
  }
}

// Value noise: lattice values in [0,1) hashed from (seed, octave, x, y),
// smoothstep-weighted bilinear interpolation.
inline double lattice(uint64_t seed, int oct, int64_t x, int64_t y) {
  const uint64_t key = ((uint64_t)(uint32_t)x << 32) ^ (uint64_t)(uint32_t)y;
  return u01(seed + 0x51ED270B27A1C3ull * (uint64_t)(oct + 1), key, 0);
}

inline double value_noise(uint64_t seed, int oct, double x, double y) {
  const double fx = std::floor(x), fy = std::floor(y);
  const int64_t ix = (int64_t)fx, iy = (int64_t)fy;
  double tx = x - fx, ty = y - fy;
  tx = tx * tx * (3.0 - 2.0 * tx);
  ty = ty * ty * (3.0 - 2.0 * ty);
  const double a = lattice(seed, oct, ix, iy), b = lattice(seed, oct, ix + 1, iy);
  const double c = lattice(seed, oct, ix, iy + 1),
               d = lattice(seed, oct, ix + 1, iy + 1);
  const double ab = a + (b - a) * tx, cd = c + (d - c) * tx;
  return ab + (cd - ab) * ty;
}

}  // namespace

extern "C" {

void lbvh_wl_uniform(uint64_t seed, uint32_t n, double h, float *tri) {
  parallel_for(n, [=](uint64_t i) {
    double c[3], v[9];
    for (int k = 0; k < 3; ++k) c[k] = u01(seed, i, k);
    for (int j = 0; j < 9; ++j)
      v[j] = c[j % 3] + (2.0 * u01(seed, i, 3 + j) - 1.0) * h;
    put_tri(tri + 12 * i, v);
  });
}

void lbvh_wl_clustered(uint64_t seed, uint32_t n, double dup, float *tri) {
  const uint64_t mseed = seed ^ 0xC1A5C1A5ull;
  double means[64][3];
  for (int m = 0; m < 64; ++m)
    for (int k = 0; k < 3; ++k) means[m][k] = u01(mseed, m, k);
  parallel_for(n, [&](uint64_t i) {
    const int m = std::min(63, (int)(u01(seed, i, 0) * 64.0));
    double c[3], v[9];
    for (int k = 0; k < 3; ++k) {
      const double g = gauss(seed, i, 1 + 2 * k);
      c[k] = std::min(1.0, std::max(0.0, means[m][k] + 0.02 * g));
    }
    for (int j = 0; j < 9; ++j)
      v[j] = c[j % 3] + (2.0 * u01(seed, i, 8 + j) - 1.0) * 0.005;
    put_tri(tri + 12 * i, v);
  });
  // Duplicates in increasing order, so a copy of a copy sees its final donor.
  for (uint64_t i = 1; i < n; ++i) {
    if (u01(seed, i, 20) < dup) {
      const uint64_t donor = std::min<uint64_t>(i - 1, (uint64_t)(u01(seed, i, 21) * i));
      std::copy(tri + 12 * donor, tri + 12 * donor + 12, tri + 12 * i);
    }
  }
}

void lbvh_wl_heightfield(uint64_t seed, uint32_t nx, uint32_t ny, float *tri) {
  const uint64_t vx = (uint64_t)nx + 1, vy = (uint64_t)ny + 1;
  std::vector<float> z(vx * vy);
  parallel_for(vx * vy, [&](uint64_t k) {
    const uint64_t ix = k % vx, iy = k / vx;
    const double x = (double)ix / nx, y = 0.5 * (double)iy / ny;
    double h = 0.0, amp = 1.0, freq = 8.0, norm = 0.0;
    for (int o = 0; o < 4; ++o) {
      h += amp * value_noise(seed, o, x * freq, y * freq);
      norm += amp;
      amp *= 0.5;
      freq *= 2.0;
    }
    z[k] = (float)(0.1 * h / norm);
  });
  parallel_for((uint64_t)nx * ny, [&](uint64_t cell) {
    const uint64_t cx = cell % nx, cy = cell / nx;
    auto vert = [&](uint64_t ix, uint64_t iy, double *out) {
      out[0] = (double)ix / nx;
      out[1] = 0.5 * (double)iy / ny;
      out[2] = z[iy * vx + ix];
    };
    double v00[3], v10[3], v01[3], v11[3];
    vert(cx, cy, v00); vert(cx + 1, cy, v10);
    vert(cx, cy + 1, v01); vert(cx + 1, cy + 1, v11);
    double a[9] = {v00[0], v00[1], v00[2], v10[0], v10[1], v10[2],
                   v11[0], v11[1], v11[2]};
    double b[9] = {v00[0], v00[1], v00[2], v11[0], v11[1], v11[2],
                   v01[0], v01[1], v01[2]};
    put_tri(tri + 24 * cell, a);
    put_tri(tri + 24 * cell + 12, b);
  });
}

void lbvh_wl_log_uniform_sizes(uint64_t seed, uint32_t count, uint32_t lo,
                               uint32_t hi, uint32_t *sizes) {
  const double a = std::log((double)lo), b = std::log((double)hi);
  for (uint32_t k = 0; k < count; ++k) {
    double s = std::floor(std::exp(a + (b - a) * u01(seed, k, 0)));
    sizes[k] = (uint32_t)std::min<double>(hi, std::max<double>(lo, s));
  }
}

void lbvh_wl_jitter(uint64_t seed, uint32_t n, double sigma, const float *in,
                    float *out) {
  parallel_for(n, [=](uint64_t i) {
    for (int k = 0; k < 3; ++k) {
      for (int c = 0; c < 3; ++c)
        out[12 * i + 4 * k + c] =
            (float)((double)in[12 * i + 4 * k + c] +
                    sigma * gauss(seed, i, 2 * (3 * k + c)));
      out[12 * i + 4 * k + 3] = in[12 * i + 4 * k + 3];
    }
  });
}

void lbvh_wl_bounds(const float *tri, uint32_t n, float *out) {
  double mn[3] = {1e30, 1e30, 1e30}, mx[3] = {-1e30, -1e30, -1e30};
  for (uint64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) {
        const double x = tri[12 * i + 4 * k + c];
        mn[c] = std::min(mn[c], x);
        mx[c] = std::max(mx[c], x);
      }
  for (int c = 0; c < 3; ++c) {
    out[c] = (float)mn[c];
    out[3 + c] = (float)mx[c];
  }
}

void lbvh_wl_uniform01(uint64_t seed, uint64_t first, uint32_t count,
                       double *out) {
  for (uint32_t k = 0; k < count; ++k) out[k] = u01(seed, first + k, 0);
}

}  // extern "C"
