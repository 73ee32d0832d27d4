// scene.cpp — synthetic multi-view scenes with analytic depth.
//
// Renders gray images of simple textured geometry (spheres, planes, boxes)
// seen by pinhole cameras, plus the exact axial depth of every pixel centre.
// The texture is band-limited 3D gradient noise evaluated at the world-space
// hit point, so one surface point has the same value in every view; pixels
// are box-filtered over ss x ss sub-samples (camera integration).
//
// It also restates the reference's own unit-test scene (tests/unit/test_mvs.cpp:
// 56-204: plane at z = 0.35 m, TextureValue of the world coordinates, toe-in
// cameras) so that test can be replayed through the GPU path.
//
// Input generation only: it is not part of the PatchMatch hot path.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "recon_mvs_hip.h"

extern "C" {

// Scene description.  Primitives are in world coordinates (metres).
typedef struct recon_scene_desc {
  int32_t num_spheres;
  int32_t num_planes;
  int32_t num_boxes;
  int32_t supersample;      // sub-samples per pixel per axis (>= 1)
  double spheres[8][4];     // centre xyz, radius
  double planes[8][6];      // point xyz, normal xyz
  double boxes[8][6];       // min xyz, max xyz (axis aligned)
  double texture_freq;      // base frequency of the noise, cycles per metre
  int32_t octaves;          // noise octaves (frequency doubles, amplitude halves)
  uint32_t seed;            // texture seed
  double contrast;          // value = 0.5 + contrast * noise
  double noise_sigma;       // additive Gaussian pixel noise (value units)
  uint32_t noise_seed;      // per-image seed of the pixel noise
  int32_t reserved;
} recon_scene_desc;

}  // extern "C"

namespace {

inline uint32_t Mix(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

inline uint32_t LatticeHash(int32_t x, int32_t y, int32_t z, uint32_t seed) {
  uint32_t h = Mix(seed ^ 0x9e3779b9U);
  h = Mix(h ^ static_cast<uint32_t>(x));
  h = Mix(h ^ static_cast<uint32_t>(y));
  h = Mix(h ^ static_cast<uint32_t>(z));
  return h;
}

// Gradient at a lattice point: one of 12 cube-edge directions.
inline double Grad(uint32_t h, double x, double y, double z) {
  switch (h % 12u) {
    case 0: return x + y;
    case 1: return -x + y;
    case 2: return x - y;
    case 3: return -x - y;
    case 4: return x + z;
    case 5: return -x + z;
    case 6: return x - z;
    case 7: return -x - z;
    case 8: return y + z;
    case 9: return -y + z;
    case 10: return y - z;
    default: return -y - z;
  }
}

inline double Fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
inline double Lerp(double a, double b, double t) { return a + t * (b - a); }

double GradientNoise(double x, double y, double z, uint32_t seed) {
  const double fx = std::floor(x), fy = std::floor(y), fz = std::floor(z);
  const int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy),
                iz = static_cast<int32_t>(fz);
  const double dx = x - fx, dy = y - fy, dz = z - fz;
  const double u = Fade(dx), v = Fade(dy), w = Fade(dz);
  double c[2][2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int k = 0; k < 2; ++k)
        c[a][b][k] = Grad(LatticeHash(ix + a, iy + b, iz + k, seed), dx - a, dy - b, dz - k);
  const double x00 = Lerp(c[0][0][0], c[1][0][0], u);
  const double x10 = Lerp(c[0][1][0], c[1][1][0], u);
  const double x01 = Lerp(c[0][0][1], c[1][0][1], u);
  const double x11 = Lerp(c[0][1][1], c[1][1][1], u);
  return Lerp(Lerp(x00, x10, v), Lerp(x01, x11, v), w);
}

double Texture(const recon_scene_desc& s, const double p[3]) {
  double sum = 0.0, amp = 1.0, freq = s.texture_freq, norm = 0.0;
  for (int o = 0; o < s.octaves; ++o) {
    sum += amp * GradientNoise(p[0] * freq, p[1] * freq, p[2] * freq, s.seed + 101u * o);
    norm += amp;
    amp *= 0.5;
    freq *= 2.0;
  }
  const double v = 0.5 + s.contrast * sum / (norm > 0.0 ? norm : 1.0);
  return std::min(1.0, std::max(0.0, v));
}

// Nearest positive ray hit of all primitives; returns t (or -1).
double Intersect(const recon_scene_desc& s, const double o[3], const double d[3]) {
  double best = -1.0;
  auto take = [&](double t) {
    if (t > 1e-9 && (best < 0.0 || t < best)) best = t;
  };
  for (int i = 0; i < s.num_spheres; ++i) {
    const double* S = s.spheres[i];
    const double oc[3] = {o[0] - S[0], o[1] - S[1], o[2] - S[2]};
    const double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
    const double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - S[3] * S[3];
    const double disc = b * b - a * c;
    if (disc < 0.0) continue;
    const double sq = std::sqrt(disc);
    const double t0 = (-b - sq) / a, t1 = (-b + sq) / a;
    if (t0 > 1e-9) take(t0);
    else take(t1);
  }
  for (int i = 0; i < s.num_planes; ++i) {
    const double* P = s.planes[i];
    const double den = P[3] * d[0] + P[4] * d[1] + P[5] * d[2];
    if (std::fabs(den) < 1e-12) continue;
    const double t = (P[3] * (P[0] - o[0]) + P[4] * (P[1] - o[1]) + P[5] * (P[2] - o[2])) / den;
    take(t);
  }
  for (int i = 0; i < s.num_boxes; ++i) {
    const double* B = s.boxes[i];
    double t0 = -1e300, t1 = 1e300;
    bool miss = false;
    for (int k = 0; k < 3; ++k) {
      if (std::fabs(d[k]) < 1e-15) {
        if (o[k] < B[k] || o[k] > B[k + 3]) miss = true;
        continue;
      }
      double a = (B[k] - o[k]) / d[k], b = (B[k + 3] - o[k]) / d[k];
      if (a > b) std::swap(a, b);
      t0 = std::max(t0, a);
      t1 = std::min(t1, b);
    }
    if (miss || t0 > t1) continue;
    if (t0 > 1e-9) take(t0);
    else take(t1);
  }
  return best;
}

inline double HashUniform(uint32_t a, uint32_t b, uint32_t c) {
  const uint32_t h = Mix(Mix(Mix(a ^ 0x85ebca6bU) ^ b) ^ c);
  return (static_cast<double>(h) + 0.5) / 4294967296.0;
}

// tests/unit/test_mvs.cpp:92-104 (TextureValue), restated.
inline uint8_t RefTextureValue(int x, int y) {
  const double u = static_cast<double>(x) * 0.08;
  const double v = static_cast<double>(y) * 0.11;
  double val = 128.0 + 60.0 * std::sin(u) * std::cos(v * 1.3) +
               30.0 * std::sin(u * 2.7 + v * 0.4) + 18.0 * std::cos(u * 4.1 - v * 3.3);
  const uint32_t h = static_cast<uint32_t>(x * 73856093u ^ y * 19349663u);
  val += static_cast<double>(h & 0xFF) * 0.15 - 19.0;
  return static_cast<uint8_t>(std::clamp(val, 0.0, 255.0));
}

void MatTVec(const double R[9], const double v[3], double o[3]) {
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    s += R[0 * 3 + i] * v[0];
    s += R[1 * 3 + i] * v[1];
    s += R[2 * 3 + i] * v[2];
    o[i] = s;
  }
}

}  // namespace

extern "C" {

// Renders one view.  image: float [h][w] in [0,1]; depth: float [h][w] axial
// depth of the pixel-centre ray (0 where the ray hits nothing); either may be
// NULL.  Pinhole only (distortion coefficients are ignored).
RECON_MVS_API int32_t recon_scene_render(const recon_scene_desc* s,
                                         const recon_mvs_camera* cam,
                                         float* image, float* depth) {
  if (!s || !cam || cam->width <= 0 || cam->height <= 0) return RECON_MVS_ERR_INVALID_ARGUMENT;
  const int w = cam->width, h = cam->height;
  const int ss = std::max(1, s->supersample);
  double origin[3];
  {
    double q[3];
    MatTVec(cam->R, cam->t, q);
    origin[0] = -q[0];
    origin[1] = -q[1];
    origin[2] = -q[2];
  }
#pragma omp parallel for schedule(dynamic, 4)
  for (int y = 0; y < h; ++y) {
    for (int x = 0; x < w; ++x) {
      const size_t idx = static_cast<size_t>(y) * w + x;
      if (depth) {
        const double rc[3] = {(x + 0.5 - cam->cx) / cam->fx, (y + 0.5 - cam->cy) / cam->fy, 1.0};
        double d[3];
        MatTVec(cam->R, rc, d);
        const double t = Intersect(*s, origin, d);
        depth[idx] = t > 0.0 ? static_cast<float>(t) : 0.0f;  // rc.z = 1 -> t = axial depth
      }
      if (image) {
        double acc = 0.0;
        int hits = 0;
        for (int j = 0; j < ss; ++j) {
          for (int i = 0; i < ss; ++i) {
            const double px = x + (i + 0.5) / ss, py = y + (j + 0.5) / ss;
            const double rc[3] = {(px - cam->cx) / cam->fx, (py - cam->cy) / cam->fy, 1.0};
            double d[3];
            MatTVec(cam->R, rc, d);
            const double t = Intersect(*s, origin, d);
            if (t <= 0.0) continue;
            const double p[3] = {origin[0] + t * d[0], origin[1] + t * d[1], origin[2] + t * d[2]};
            acc += Texture(*s, p);
            ++hits;
          }
        }
        double v = hits ? acc / (ss * ss) : 0.0;
        if (s->noise_sigma > 0.0) {
          const uint32_t a = static_cast<uint32_t>(idx);
          const double u1 = HashUniform(s->noise_seed, a, 1u);
          const double u2 = HashUniform(s->noise_seed, a, 2u);
          v += s->noise_sigma * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
        }
        image[idx] = static_cast<float>(v);
      }
    }
  }
  return RECON_MVS_OK;
}

// The reference unit-test scene (tests/unit/test_mvs.cpp:112-204):
// 160x120, f = 400, plane z = 0.35 m, num_views cameras on x in
// [-baseline, +baseline] toed in at (0, 0, 0.35).  Writes num_views cameras,
// num_views uint8 images and the reference view's truth depth.
RECON_MVS_API int32_t recon_scene_reference_plane(int32_t num_views, double baseline_m,
                                                  recon_mvs_camera* cams, uint8_t* images,
                                                  float* truth_depth) {
  if (num_views < 1) return RECON_MVS_ERR_INVALID_ARGUMENT;
  const int W = 160, H = 120;
  const double plane_depth = 0.35;
  for (int v = 0; v < num_views; ++v) {
    recon_mvs_camera& c = cams[v];
    std::memset(&c, 0, sizeof(c));
    c.fx = 400.0; c.fy = 400.0; c.cx = 80.0; c.cy = 60.0;
    c.width = W; c.height = H;
    c.R[0] = c.R[4] = c.R[8] = 1.0;
    if (v == 0) continue;
    const double frac = num_views > 1 ? static_cast<double>(v) / (num_views - 1) : 0.0;
    const double dx = (frac * 2.0 - 1.0) * baseline_m;
    // forward = normalized(p_ref - center)
    double f[3] = {0.0 - dx, 0.0, plane_depth - 0.0};
    {
      double s = 0.0;
      s += f[0] * f[0]; s += f[1] * f[1]; s += f[2] * f[2];
      const double n = std::sqrt(s);
      for (double& q : f) q = n < 1e-15 ? 0.0 : q / n;
    }
    const double down_w[3] = {0.0, -1.0, 0.0};
    double r[3] = {f[1] * down_w[2] - f[2] * down_w[1], f[2] * down_w[0] - f[0] * down_w[2],
                   f[0] * down_w[1] - f[1] * down_w[0]};
    {
      double s = 0.0;
      s += r[0] * r[0]; s += r[1] * r[1]; s += r[2] * r[2];
      double n = std::sqrt(s);
      if (n < 1e-9) { r[0] = 1.0; r[1] = 0.0; r[2] = 0.0; }
      s = 0.0;
      s += r[0] * r[0]; s += r[1] * r[1]; s += r[2] * r[2];
      n = std::sqrt(s);
      if (!(n < 1e-15)) for (double& q : r) q /= n;
    }
    const double dn[3] = {f[1] * r[2] - f[2] * r[1], f[2] * r[0] - f[0] * r[2],
                          f[0] * r[1] - f[1] * r[0]};
    for (int j = 0; j < 3; ++j) {
      c.R[0 * 3 + j] = r[j];
      c.R[1 * 3 + j] = dn[j];
      c.R[2 * 3 + j] = f[j];
    }
    const double center[3] = {dx, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      s += c.R[i * 3 + 0] * center[0];
      s += c.R[i * 3 + 1] * center[1];
      s += c.R[i * 3 + 2] * center[2];
      c.t[i] = -s;
    }
  }
  for (int v = 0; v < num_views; ++v) {
    const recon_mvs_camera& c = cams[v];
    uint8_t* img = images + static_cast<size_t>(v) * W * H;
    std::memset(img, 0, static_cast<size_t>(W) * H);
    double origin[3];
    {
      double q[3];
      MatTVec(c.R, c.t, q);
      origin[0] = -q[0]; origin[1] = -q[1]; origin[2] = -q[2];
    }
    for (int y = 0; y < H; ++y) {
      for (int x = 0; x < W; ++x) {
        const double ray[3] = {(x + 0.5 - c.cx) / c.fx, (y + 0.5 - c.cy) / c.fy, 1.0};
        double d[3];
        MatTVec(c.R, ray, d);
        if (std::fabs(d[2]) < 1e-9) continue;
        const double t = (plane_depth - origin[2]) / d[2];
        if (t <= 0.0) continue;
        const double p[3] = {origin[0] + d[0] * t, origin[1] + d[1] * t, origin[2] + d[2] * t};
        if (v == 0 && truth_depth) truth_depth[static_cast<size_t>(y) * W + x] = static_cast<float>(p[2]);
        const int tx = static_cast<int>(std::lround(p[0] * 1500.0)) + 512;
        const int ty = static_cast<int>(std::lround(p[1] * 1500.0)) + 512;
        img[static_cast<size_t>(y) * W + x] = RefTextureValue(tx, ty);
      }
    }
  }
  return RECON_MVS_OK;
}

}  // extern "C"
