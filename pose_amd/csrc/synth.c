/*
 * Seeded synthetic rectified stereo pairs for the benchmark workloads.
 *
 * Scene: band-limited 4-octave gradient (Perlin) noise texture; 3-6 slanted
 * disparity planes on a Voronoi partition; right image = left texture sampled
 * at x + d (linear along x); Gaussian noise sigma = 2 (Box-Muller); clamp to
 * [0, 255].  With gain != 0 the right image is multiplied by a smooth gain
 * field in [0.8, 1.2].  Deterministic for a seed; rows run in parallel with
 * OpenMP (every pixel's noise comes from a counter-based hash, so the output
 * does not depend on the thread count).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

static uint64_t mix64(uint64_t z) {
  /* splitmix64 finaliser */
  z += 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static double unit(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

static float fade(float t) { return t * t * t * (t * (t * 6.f - 15.f) + 10.f); }

/* gradient noise: lattice gradients from a hash of (octave, ix, iy, seed) */
static float grad_noise(float x, float y, int octave, uint64_t seed) {
  const int ix = (int)floorf(x), iy = (int)floorf(y);
  const float fx = x - ix, fy = y - iy;
  float v[4];
  for (int c = 0; c < 4; c++) {
    const int dx = c & 1, dy = c >> 1;
    const uint64_t h = mix64(seed ^ ((uint64_t)(uint32_t)(ix + dx) << 32) ^
                             (uint64_t)(uint32_t)(iy + dy) ^ ((uint64_t)octave << 56));
    const float a = (float)(unit(h) * 6.283185307179586);
    v[c] = cosf(a) * (fx - dx) + sinf(a) * (fy - dy);
  }
  const float u = fade(fx), w = fade(fy);
  const float top = v[0] + (v[1] - v[0]) * u, bot = v[2] + (v[3] - v[2]) * u;
  return (top + (bot - top) * w) * 1.41f;
}

static float texture_at(float x, float y, uint64_t seed) {
  float t = 0.f, amp = 1.f, total = 0.f, spacing = 16.f;
  for (int o = 0; o < 4; o++) {
    t += amp * grad_noise(x / spacing, y / spacing, o, seed);
    total += amp;
    amp *= 0.5f;
    spacing *= 0.5f;
  }
  return 128.f + 160.f * (t / total);
}

static float gauss(uint64_t h1, uint64_t h2) {
  const double u1 = unit(h1) + 1e-300, u2 = unit(h2);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

static uint8_t to_u8(float v) {
  v = v < 0.f ? 0.f : v > 255.f ? 255.f : v;
  return (uint8_t)lrintf(v);
}

/* left, right: w*h bytes; disp (may be NULL): w*h floats of left-view truth. */
int sgm_synth_pair(int w, int h, int disparities, int min_disparity, uint64_t seed, float noise,
                   int gain, uint8_t *left, uint8_t *right, float *disp) {
  if (w < 1 || h < 1 || disparities < 1) return -1;
  uint64_t s = mix64(seed * 0x2545F4914F6CDD1DULL + 17);
  const int count = 3 + (int)(unit(s = mix64(s)) * 4); /* 3..6 planes */
  float lo = min_disparity + 2.f, hi = min_disparity + disparities - 3.f;
  if (hi < lo) {
    lo = (float)min_disparity;
    hi = (float)(min_disparity + disparities - 1);
  }
  float sx[8], sy[8], base[8], gx[8], gy[8];
  const float span = (hi - lo) * 0.15f;
  for (int i = 0; i < count; i++) {
    sx[i] = (float)(unit(s = mix64(s)) * w);
    sy[i] = (float)(unit(s = mix64(s)) * h);
    base[i] = lo + 0.15f * (hi - lo) + (float)unit(s = mix64(s)) * 0.7f * (hi - lo);
    gx[i] = ((float)unit(s = mix64(s)) * 2.f - 1.f) * span / (float)w;
    gy[i] = ((float)unit(s = mix64(s)) * 2.f - 1.f) * span / (float)h;
  }
  const uint64_t tex_seed = mix64(s ^ 0xA5A5A5A5ULL), noise_seed = mix64(s ^ 0x5A5A5A5AULL);
#pragma omp parallel for schedule(static)
  for (int y = 0; y < h; y++) {
    for (int x = 0; x < w; x++) {
      /* plane of the nearest seed */
      int best = 0;
      float bd = 1e30f;
      for (int i = 0; i < count; i++) {
        const float dx = x - sx[i], dy = y - sy[i], d2 = dx * dx + dy * dy;
        if (d2 < bd) {
          bd = d2;
          best = i;
        }
      }
      float d = base[best] + gx[best] * (x - sx[best]) + gy[best] * (y - sy[best]);
      d = d < lo ? lo : d > hi ? hi : d;
      const size_t p = (size_t)y * w + x;
      if (disp) disp[p] = d;
      float l = texture_at((float)x, (float)y, tex_seed);
      /* right(x) = texture(x + d), linear along x between integer samples */
      const float xs = x + d;
      const float x0 = floorf(xs), f = xs - x0;
      float r = texture_at(x0, (float)y, tex_seed) * (1.f - f) +
                texture_at(x0 + 1.f, (float)y, tex_seed) * f;
      if (gain) {
        float g = grad_noise(x / (float)(w > h ? h : w) * 4.f, y / (float)(w > h ? h : w) * 4.f,
                             9, tex_seed ^ 0x77ULL);
        g = g < -1.f ? -1.f : g > 1.f ? 1.f : g;
        r *= 1.f + 0.2f * g;
      }
      if (noise > 0.f) {
        const uint64_t k = noise_seed ^ (p * 4);
        l += noise * gauss(mix64(k), mix64(k + 1));
        r += noise * gauss(mix64(k + 2), mix64(k + 3));
      }
      left[p] = to_u8(l);
      right[p] = to_u8(r);
    }
  }
  return 0;
}
