/*
 * pm_oracle.c — TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * Plain-C restatement of the reference CPU PatchMatch MVS path
 * (core/src/mvs/patchmatch_stereo.cpp) used as the checker of the GPU
 * library's reference mode.  Every function cites the reference lines it
 * follows and performs the same double operations in the same order, so a
 * build without FMA contraction reproduces the reference bit for bit.  It is
 * never linked into, loaded by or called from the product path.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_xorshift_*        published xorshift64* (Vigna 2016) known answers
 *   orc_sample_bilinear   closed forms (exact on affine ramps, pixel centres)
 *   orc_compute_ncc       the compiled reference (oracle/_ref) + invariants
 *   orc_evaluate          the compiled reference through PatchMatchStereo
 *                         with iterations = 0 (initial costs)
 *   orc_patchmatch        the compiled reference, bit-exact, several
 *                         num_threads values, golden fixtures
 *   orc_refine_subpixel   the compiled reference, bit-exact
 *   orc_fuse              the compiled reference (key-sorted sets)
 *   orc_select_views      the compiled reference
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pm_oracle.h"

/* ------------------------------------------------------------------------- */
/* Camera: RemoveDistortion (camera.cpp:47-129)                              */
/* ------------------------------------------------------------------------- */

void orc_remove_distortion(const recon_mvs_camera* c, double x_d, double y_d,
                           double* out_x, double* out_y) {
  if (c->k1 == 0.0 && c->k2 == 0.0 && c->k3 == 0.0 && c->p1 == 0.0 &&
      c->p2 == 0.0) {
    *out_x = x_d;
    *out_y = y_d;
    return;
  }
  double x = x_d, y = y_d;
  for (int iter = 0; iter < 50; ++iter) {
    double x_r, y_r;
    orc_apply_distortion(c, x, y, &x_r, &y_r);
    const double ex = x_d - x_r;
    const double ey = y_d - y_r;
    const double err = sqrt(ex * ex + ey * ey);
    if (err < 1e-15) break;
    const double r2 = x * x + y * y;
    const double r4 = r2 * r2;
    const double radial = 1.0 + c->k1 * r2 + c->k2 * r4 + c->k3 * r4 * r2;
    const double dr = c->k1 + 2.0 * c->k2 * r2 + 3.0 * c->k3 * r4;
    const double j00 = radial + 2.0 * x * x * dr + 2.0 * c->p1 * y +
                       c->p2 * (3.0 * x * x + y * y);
    const double j01 = 2.0 * x * y * dr + c->p1 * (x * x + 3.0 * y * y) +
                       2.0 * c->p2 * x * y;
    const double j10 = 2.0 * x * y * dr + 2.0 * c->p1 * x +
                       c->p2 * (x * x + 3.0 * y * y);
    const double j11 = radial + 2.0 * y * y * dr +
                       c->p1 * (3.0 * x * x + y * y) + 2.0 * c->p2 * x * y;
    const double det = j00 * j11 - j01 * j10;
    if (fabs(det) < 1e-12) break;
    const double dx = (j11 * ex - j01 * ey) / det;
    const double dy = (-j10 * ex + j00 * ey) / det;
    double lambda = 1.0;
    int accepted = 0;
    for (int ls = 0; ls < 20; ++ls) {
      const double xn = x + lambda * dx;
      const double yn = y + lambda * dy;
      double x_rn, y_rn;
      orc_apply_distortion(c, xn, yn, &x_rn, &y_rn);
      const double rn = sqrt((x_d - x_rn) * (x_d - x_rn) +
                             (y_d - y_rn) * (y_d - y_rn));
      if (rn < err) {
        x = xn;
        y = yn;
        accepted = 1;
        break;
      }
      lambda *= 0.5;
    }
    if (!accepted) break;
  }
  *out_x = x;
  *out_y = y;
}

void orc_pixel_to_ray_batch(const recon_mvs_camera* c, int n, const double* uv,
                            double* rays) {
  for (int i = 0; i < n; ++i) orc_pixel_to_ray(c, uv[2 * i], uv[2 * i + 1], rays + 3 * i);
}

void orc_project_batch(const recon_mvs_camera* c, int n, const double* p,
                       double* uv, int32_t* ok) {
  for (int i = 0; i < n; ++i) {
    double u = 0.0, v = 0.0;
    ok[i] = orc_project(c, p + 3 * i, &u, &v);
    uv[2 * i] = u;
    uv[2 * i + 1] = v;
  }
}

/* ------------------------------------------------------------------------- */
/* xorshift64* (patchmatch_stereo.cpp:104-125)                               */
/* ------------------------------------------------------------------------- */

static inline uint32_t xs_next(uint64_t* s) {
  *s ^= *s >> 12;
  *s ^= *s << 25;
  *s ^= *s >> 27;
  return (uint32_t)((*s * 0x2545F4914F6CDD1DULL) >> 32);
}
static inline uint64_t xs_seed(uint64_t seed) {
  return seed ? seed : 0x9E3779B97F4A7C15ULL;
}
static inline double xs_uniform(uint64_t* s) {
  return (double)xs_next(s) / 4294967296.0;
}
static inline double xs_signed(uint64_t* s) { return xs_uniform(s) * 2.0 - 1.0; }

void orc_xorshift_stream(uint64_t seed, int n, uint32_t* out) {
  uint64_t s = xs_seed(seed);
  for (int i = 0; i < n; ++i) out[i] = xs_next(&s);
}

/* ------------------------------------------------------------------------- */
/* SampleBilinear (patchmatch_stereo.cpp:78-100)                             */
/* ------------------------------------------------------------------------- */

static inline int sample_bilinear(const uint8_t* img, int w, int h, double u,
                                  double v, double* out) {
  if (u < 0.0 || v < 0.0 || u > (double)(w - 1) || v > (double)(h - 1)) return 0;
  const int x0 = (int)floor(u);
  const int y0 = (int)floor(v);
  const int x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1;
  const int y1 = y0 + 1 < h - 1 ? y0 + 1 : h - 1;
  const double fx = u - x0;
  const double fy = v - y0;
  const double p00 = img[y0 * w + x0];
  const double p10 = img[y0 * w + x1];
  const double p01 = img[y1 * w + x0];
  const double p11 = img[y1 * w + x1];
  const double top = p00 * (1.0 - fx) + p10 * fx;
  const double bot = p01 * (1.0 - fx) + p11 * fx;
  *out = top * (1.0 - fy) + bot * fy;
  return 1;
}

int orc_sample_bilinear(const uint8_t* img, int w, int h, double u, double v,
                        double* out) {
  return sample_bilinear(img, w, h, u, v, out);
}

/* ------------------------------------------------------------------------- */
/* ComputeNcc (patchmatch_stereo.cpp:214-274)                                */
/* ------------------------------------------------------------------------- */

static recon_mvs_ncc_result compute_ncc(const uint8_t* ref, int rw, int rh,
                                        const uint8_t* src, int sw, int sh,
                                        double ru, double rv, double su,
                                        double sv, int patch) {
  recon_mvs_ncc_result r;
  memset(&r, 0, sizeof(r));
  const int half = patch / 2;
  double s_r = 0.0, s_s = 0.0, s_r2 = 0.0, s_s2 = 0.0, s_x = 0.0;
  int valid = 0;
  for (int dy = -half; dy <= half; ++dy) {
    for (int dx = -half; dx <= half; ++dx) {
      double vr = 0.0, vs = 0.0;
      if (!sample_bilinear(ref, rw, rh, ru + dx, rv + dy, &vr)) continue;
      if (!sample_bilinear(src, sw, sh, su + dx, sv + dy, &vs)) continue;
      s_r += vr;
      s_s += vs;
      s_r2 += vr * vr;
      s_s2 += vs * vs;
      s_x += vr * vs;
      ++valid;
    }
  }
  if (valid < patch * patch / 2) return r;
  const double n = (double)valid;
  const double mr = s_r / n;
  const double ms = s_s / n;
  const double var_r = s_r2 / n - mr * mr;
  const double var_s = s_s2 / n - ms * ms;
  r.weak_texture = var_r < 12.0;
  const double denom = sqrt(var_r * var_s);
  if (denom < 1e-9) return r;
  double ncc = (s_x / n - mr * ms) / denom;
  if (ncc < -1.0) ncc = -1.0;
  if (ncc > 1.0) ncc = 1.0;
  r.ncc = ncc;
  r.valid = 1;
  r.valid_samples = valid;
  return r;
}

void orc_compute_ncc(const uint8_t* a, int aw, int ah, const uint8_t* b, int bw,
                     int bh, int n, const double* coords, int patch,
                     recon_mvs_ncc_result* out) {
  for (int i = 0; i < n; ++i) {
    const double* c = coords + 4 * i;
    out[i] = compute_ncc(a, aw, ah, b, bw, bh, c[0], c[1], c[2], c[3], patch);
  }
}

/* ------------------------------------------------------------------------- */
/* EvaluateCandidate (patchmatch_stereo.cpp:344-418)                         */
/* ------------------------------------------------------------------------- */

typedef struct {
  const uint8_t* ref;
  int rw, rh;
  int nsrc;
  const uint8_t* const* src;
  const int32_t* sw;
  const int32_t* sh;
  const recon_mvs_camera* ref_cam;
  const recon_mvs_camera* src_cams;
  const recon_mvs_config* cfg;
} orc_problem;

static recon_mvs_cost_result evaluate(const orc_problem* p, double u, double v,
                                      double inv) {
  recon_mvs_cost_result r;
  r.cost = 1e30;
  r.ncc_sum = 0.0;
  r.weight_sum = 0.0;
  r.weak_texture_count = 0;
  r.valid_views = 0;
  if (inv <= 1e-9) return r;
  const double depth = 1.0 / inv; /* InverseDepthToDepth, inv > 1e-9 here */
  if (depth < p->cfg->min_depth_m || depth > p->cfg->max_depth_m) return r;
  double ray[3];
  orc_pixel_to_ray(p->ref_cam, u, v, ray);
  const double pc[3] = {ray[0] * depth, ray[1] * depth, ray[2] * depth};
  double pw[3];
  orc_cam_to_world(p->ref_cam, pc, pw);
  for (int k = 0; k < p->nsrc; ++k) {
    double ps[3];
    orc_world_to_cam(&p->src_cams[k], pw, ps);
    if (ps[2] <= 1e-6) continue;
    double su = 0.0, sv = 0.0;
    if (!orc_project(&p->src_cams[k], ps, &su, &sv)) continue;
    const recon_mvs_ncc_result n =
        compute_ncc(p->ref, p->rw, p->rh, p->src[k], p->sw[k], p->sh[k], u, v,
                    su, sv, p->cfg->patch_size);
    if (!n.valid) continue;
    ++r.valid_views;
    if (n.weak_texture) ++r.weak_texture_count;
    r.ncc_sum += 1.0 * n.ncc;
    r.weight_sum += 1.0;
  }
  if (r.weight_sum <= 0.0) return r;
  const double mean = r.ncc_sum / r.weight_sum;
  r.cost = 1.0 - mean;
  if (r.cost > 3.0 * p->cfg->photo_threshold) r.cost = 1e30;
  return r;
}

static orc_problem make_problem(const uint8_t* ref, int rw, int rh, int nsrc,
                                const uint8_t* const* src, const int32_t* sw,
                                const int32_t* sh,
                                const recon_mvs_camera* ref_cam,
                                const recon_mvs_camera* src_cams,
                                const recon_mvs_config* cfg) {
  orc_problem p = {ref, rw, rh, nsrc, src, sw, sh, ref_cam, src_cams, cfg};
  return p;
}

void orc_evaluate(const uint8_t* ref, int rw, int rh, int nsrc,
                  const uint8_t* const* src, const int32_t* sw,
                  const int32_t* sh, const recon_mvs_camera* ref_cam,
                  const recon_mvs_camera* src_cams, const recon_mvs_config* cfg,
                  int n, const double* uvi, recon_mvs_cost_result* out) {
  const orc_problem p =
      make_problem(ref, rw, rh, nsrc, src, sw, sh, ref_cam, src_cams, cfg);
#pragma omp parallel for schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) out[i] = evaluate(&p, uvi[3 * i], uvi[3 * i + 1], uvi[3 * i + 2]);
}

static double mean_baseline(const orc_problem* p) {
  double rc[3];
  orc_center(p->ref_cam, rc);
  double sum = 0.0;
  for (int k = 0; k < p->nsrc; ++k) {
    double sc[3];
    orc_center(&p->src_cams[k], sc);
    const double d[3] = {sc[0] - rc[0], sc[1] - rc[1], sc[2] - rc[2]};
    sum += orc_norm3(d);
  }
  return sum / (double)p->nsrc;
}

static inline double depth_to_inv(double d) { return d < 1e-9 ? 1e9 : 1.0 / d; }
static inline double inv_to_depth(double x) { return x <= 1e-9 ? 1e9 : 1.0 / x; }

/* ------------------------------------------------------------------------- */
/* PatchMatchStereo (patchmatch_stereo.cpp:422-764)                          */
/* Argument checks of lines 434-468 are the caller's (index contract).       */
/* ------------------------------------------------------------------------- */

uint32_t orc_patchmatch(const uint8_t* ref, int rw, int rh, int nsrc,
                        const uint8_t* const* src, const int32_t* sw,
                        const int32_t* sh, const recon_mvs_camera* ref_cam,
                        const recon_mvs_camera* src_cams,
                        const recon_mvs_config* cfg, const double* seeds,
                        int num_seeds, float* out_depth, uint8_t* out_support,
                        float* out_conf) {
  const size_t P = (size_t)rw * rh;
  for (size_t i = 0; i < P; ++i) {
    out_depth[i] = NAN;
    out_support[i] = 0;
    out_conf[i] = 0.0f;
  }
  if (nsrc <= 0) return 0;
  const orc_problem prob =
      make_problem(ref, rw, rh, nsrc, src, sw, sh, ref_cam, src_cams, cfg);
  if (mean_baseline(&prob) < 1e-9) return 0;
  const double inv_min = depth_to_inv(cfg->max_depth_m);
  const double inv_max = depth_to_inv(cfg->min_depth_m);
  if (!(inv_max > inv_min) || inv_min <= 0.0) return 0;

  double* inv = (double*)calloc(P, sizeof(double));
  double* cost = (double*)malloc(P * sizeof(double));
  int* support = (int*)calloc(P, sizeof(int));
  uint8_t* init = (uint8_t*)calloc(P, 1);
  for (size_t i = 0; i < P; ++i) cost[i] = 1e30;

  /* Sparse seeding: 32-px grid bucket mean (lines 499-533). */
  if (seeds && num_seeds > 0) {
    const int grid = 32;
    const int gx = (rw + grid - 1) / grid;
    const int gy = (rh + grid - 1) / grid;
    double* sum = (double*)calloc((size_t)gx * gy, sizeof(double));
    int* cnt = (int*)calloc((size_t)gx * gy, sizeof(int));
    for (int i = 0; i < num_seeds; ++i) {
      const double su = seeds[3 * i], sv = seeds[3 * i + 1], sd = seeds[3 * i + 2];
      if (su < 0 || sv < 0 || su >= rw || sv >= rh) continue;
      if (sd <= 0.0) continue;
      int cx = (int)su / grid, cy = (int)sv / grid;
      if (cx > gx - 1) cx = gx - 1;
      if (cy > gy - 1) cy = gy - 1;
      /* the reference sums each bucket in insertion order (line 526) */
      sum[(size_t)cy * gx + cx] += sd;
      cnt[(size_t)cy * gx + cx] += 1;
    }
    for (int y = 0; y < rh; ++y) {
      for (int x = 0; x < rw; ++x) {
        int cx = x / grid, cy = y / grid;
        if (cx > gx - 1) cx = gx - 1;
        if (cy > gy - 1) cy = gy - 1;
        const size_t c = (size_t)cy * gx + cx;
        if (cnt[c] == 0) continue;
        const size_t idx = (size_t)y * rw + x;
        inv[idx] = depth_to_inv(sum[c] / (double)cnt[c]);
        init[idx] = 1;
      }
    }
    free(sum);
    free(cnt);
  }

  /* Random init: one sequential stream, row-major (lines 496, 536-539). */
  {
    uint64_t s = xs_seed(0x5DEECE66DULL);
    for (size_t i = 0; i < P; ++i) {
      if (init[i]) continue;
      inv[i] = inv_min + xs_uniform(&s) * (inv_max - inv_min);
    }
  }

  /* Initial cost (lines 543-581). */
#pragma omp parallel for schedule(dynamic, 1)
  for (int y = 0; y < rh; ++y) {
    for (int x = 0; x < rw; ++x) {
      const size_t idx = (size_t)y * rw + x;
      const recon_mvs_cost_result c = evaluate(&prob, x + 0.5, y + 0.5, inv[idx]);
      cost[idx] = c.cost;
      support[idx] = c.valid_views;
    }
  }

  int threads = cfg->num_threads;
  if (threads <= 0) threads = 1; /* callers pin num_threads for parity */
  const int dirs = cfg->propagation_directions > 1 ? cfg->propagation_directions : 1;
  double* snap_inv = (double*)malloc(P * sizeof(double));
  double* snap_cost = (double*)malloc(P * sizeof(double));

  for (int iter = 0; iter < cfg->iterations; ++iter) {
    for (int dir = 0; dir < dirs; ++dir) {
      const int reverse_x = (dir == 1 || dir == 3);
      const int reverse_y = (dir >= 2);
      int block_count = rh / 16 > 1 ? rh / 16 : 1;
      if (threads < block_count) block_count = threads;
      const int rows_per_block = (rh + block_count - 1) / block_count;
      memcpy(snap_inv, inv, P * sizeof(double));
      memcpy(snap_cost, cost, P * sizeof(double));
      /* Jacobi pass: reads only the snapshot, so blocks are independent. */
#pragma omp parallel for schedule(dynamic, 1)
      for (int b = 0; b < block_count; ++b) {
        const int yb0 = b * rows_per_block;
        const int yb1 = yb0 + rows_per_block < rh ? yb0 + rows_per_block : rh;
        if (yb0 >= yb1) continue;
        uint64_t rng = xs_seed(0x5DEECE66DULL +
                               (uint64_t)b * 0x9E3779B97F4A7C15ULL +
                               (uint64_t)iter * 0xC2B2AE3D27D4EB4FULL);
        const int ys = reverse_y ? yb1 - 1 : yb0;
        const int ye = reverse_y ? yb0 - 1 : yb1;
        const int ystep = reverse_y ? -1 : 1;
        for (int y = ys; y != ye; y += ystep) {
          const int xs = reverse_x ? rw - 1 : 0;
          const int xe = reverse_x ? -1 : rw;
          const int xstep = reverse_x ? -1 : 1;
          for (int x = xs; x != xe; x += xstep) {
            const size_t idx = (size_t)y * rw + x;
            if (cfg->propagation_enabled) {
              static const int nb[4][2] = {{-1, 0}, {1, 0}, {0, -1}, {0, 1}};
              for (int k = 0; k < 4; ++k) {
                const int nx = x + nb[k][0], ny = y + nb[k][1];
                if (nx < 0 || ny < 0 || nx >= rw || ny >= rh) continue;
                const size_t nidx = (size_t)ny * rw + nx;
                if (snap_cost[nidx] >= 1e29) continue;
                const recon_mvs_cost_result c =
                    evaluate(&prob, x + 0.5, y + 0.5, snap_inv[nidx]);
                if (c.cost < cost[idx]) {
                  cost[idx] = c.cost;
                  inv[idx] = snap_inv[nidx];
                  support[idx] = c.valid_views;
                }
              }
            }
            double radius = inv[idx] * cfg->random_search_radius_ratio;
            for (int k = 0; k < cfg->random_search_count; ++k) {
              const double cand = inv[idx] + xs_signed(&rng) * radius;
              if (cand < inv_min || cand > inv_max) {
                radius *= 0.5;
                continue;
              }
              const recon_mvs_cost_result c = evaluate(&prob, x + 0.5, y + 0.5, cand);
              if (c.cost < cost[idx]) {
                cost[idx] = c.cost;
                inv[idx] = cand;
                support[idx] = c.valid_views;
              }
              radius *= 0.5;
            }
            if (xs_uniform(&rng) < cfg->restart_ratio) {
              const double cand = inv_min + xs_uniform(&rng) * (inv_max - inv_min);
              const recon_mvs_cost_result c = evaluate(&prob, x + 0.5, y + 0.5, cand);
              if (c.cost < cost[idx]) {
                cost[idx] = c.cost;
                inv[idx] = cand;
                support[idx] = c.valid_views;
              }
            }
          }
        }
      }
    }
  }

  uint32_t filled = 0;
  for (int y = 0; y < rh; ++y) {
    for (int x = 0; x < rw; ++x) {
      const size_t idx = (size_t)y * rw + x;
      if (cost[idx] >= 1e29) continue;
      const double depth = inv_to_depth(inv[idx]);
      if (depth < cfg->min_depth_m || depth > cfg->max_depth_m) continue;
      out_depth[idx] = (float)depth;
      out_support[idx] = (uint8_t)(support[idx] < 255 ? support[idx] : 255);
      double pc = cost[idx];
      if (pc < 0.0) pc = 0.0;
      if (pc > 1.0) pc = 1.0;
      out_conf[idx] = (float)(1.0 - pc);
      ++filled;
    }
  }
  free(inv);
  free(cost);
  free(support);
  free(init);
  free(snap_inv);
  free(snap_cost);
  return filled;
}

/* ------------------------------------------------------------------------- */
/* SolveLinear3x3 (patchmatch_stereo.cpp:42-75)                              */
/* ------------------------------------------------------------------------- */

static int solve3(const double A[3][3], const double b[3], double x[3]) {
  double m[3][4];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) m[i][j] = A[i][j];
    m[i][3] = b[i];
  }
  for (int col = 0; col < 3; ++col) {
    int pivot = col;
    for (int r = col + 1; r < 3; ++r)
      if (fabs(m[r][col]) > fabs(m[pivot][col])) pivot = r;
    if (fabs(m[pivot][col]) < 1e-18) return 0;
    if (pivot != col) {
      for (int j = 0; j < 4; ++j) {
        const double t = m[col][j];
        m[col][j] = m[pivot][j];
        m[pivot][j] = t;
      }
    }
    for (int r = 0; r < 3; ++r) {
      if (r == col) continue;
      const double f = m[r][col] / m[col][col];
      for (int j = col; j < 4; ++j) m[r][j] -= f * m[col][j];
    }
  }
  for (int i = 0; i < 3; ++i) {
    if (fabs(m[i][i]) < 1e-18) return 0;
    x[i] = m[i][3] / m[i][i];
  }
  return 1;
}

int orc_solve3(const double* A, const double* b, double* x) {
  const double M[3][3] = {{A[0], A[1], A[2]}, {A[3], A[4], A[5]}, {A[6], A[7], A[8]}};
  return solve3(M, b, x);
}

/* ------------------------------------------------------------------------- */
/* RefineSubpixel (patchmatch_stereo.cpp:770-979)                            */
/* ------------------------------------------------------------------------- */

uint32_t orc_refine_subpixel(const uint8_t* ref, int rw, int rh, int nsrc,
                             const uint8_t* const* src, const int32_t* sw,
                             const int32_t* sh, const recon_mvs_camera* ref_cam,
                             const recon_mvs_camera* src_cams,
                             const recon_mvs_config* cfg, float* depth,
                             float* conf) {
  if (nsrc <= 0) return 0;
  const orc_problem prob =
      make_problem(ref, rw, rh, nsrc, src, sw, sh, ref_cam, src_cams, cfg);
  const double baseline = mean_baseline(&prob);
  if (baseline < 1e-9) return 0;
  const double inv_min = depth_to_inv(cfg->max_depth_m);
  const double inv_max = depth_to_inv(cfg->min_depth_m);
  const double focal = 0.5 * (ref_cam->fx + ref_cam->fy);
  const int fit_points = cfg->subpixel_fit_points > 3 ? cfg->subpixel_fit_points : 3;
  const int half_fit = fit_points / 2;
  uint32_t refined = 0;

#pragma omp parallel for schedule(dynamic, 1) reduction(+ : refined)
  for (int y = 0; y < rh; ++y) {
    double* fit_cost = (double*)malloc((size_t)fit_points * sizeof(double));
    for (int x = 0; x < rw; ++x) {
      const size_t idx = (size_t)y * rw + x;
      const float d0 = depth[idx];
      if (!isfinite(d0)) continue;
      const double u = x + 0.5, v = y + 0.5;
      const double inv0 = depth_to_inv(d0);
      if (inv0 < inv_min || inv0 > inv_max) continue;
      const double inv_per_px = 1.0 / (focal * baseline);
      const double fit_step = inv_per_px * 0.5;
      int fit_ok = 1;
      for (int k = 0; k < fit_points; ++k) fit_cost[k] = 1e30;
      for (int k = 0; k < fit_points; ++k) {
        const double d = inv0 + (k - half_fit) * fit_step;
        if (d < inv_min || d > inv_max) {
          fit_ok = 0;
          break;
        }
        fit_cost[k] = evaluate(&prob, u, v, d).cost;
        if (fit_cost[k] >= 1e29) fit_ok = 0;
      }
      double best_d = inv0;
      double best_c = evaluate(&prob, u, v, inv0).cost;
      if (fit_ok) {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, t0 = 0, t1 = 0, t2 = 0;
        for (int k = 0; k < fit_points; ++k) {
          const double xx = (double)(k - half_fit);
          const double c = fit_cost[k];
          s0 += 1.0;
          s1 += xx;
          s2 += xx * xx;
          s3 += xx * xx * xx;
          s4 += xx * xx * xx * xx;
          t0 += c;
          t1 += xx * c;
          t2 += xx * xx * c;
        }
        const double A[3][3] = {{s4, s3, s2}, {s3, s2, s1}, {s2, s1, s0}};
        const double b[3] = {t2, t1, t0};
        double sol[3] = {0.0, 0.0, 0.0};
        if (solve3(A, b, sol)) {
          const double a = sol[0], bb = sol[1];
          if (fabs(a) > 1e-12) {
            const double vertex = -bb / (2.0 * a);
            if (vertex >= -(double)half_fit && vertex <= (double)half_fit) {
              /* The reference adds the vertex in sample-index units
                 (line 898); reproduced as is. */
              const double cand = inv0 + vertex;
              const double cc = evaluate(&prob, u, v, cand).cost;
              if (cc < best_c) {
                best_c = cc;
                best_d = cand;
              }
            }
          }
        }
      }
      {
        double lo = best_d - inv_per_px * 1.0;
        double hi = best_d + inv_per_px * 1.0;
        if (lo < inv_min) lo = inv_min;
        if (hi > inv_max) hi = inv_max;
        if (hi > lo) {
          const double inv_phi = 0.6180339887498949;
          double c1 = hi - inv_phi * (hi - lo);
          double c2 = lo + inv_phi * (hi - lo);
          double f1 = evaluate(&prob, u, v, c1).cost;
          double f2 = evaluate(&prob, u, v, c2).cost;
          for (int k = 0; k < cfg->subpixel_iterations; ++k) {
            if (f1 < f2) {
              hi = c2;
              c2 = c1;
              f2 = f1;
              c1 = hi - inv_phi * (hi - lo);
              f1 = evaluate(&prob, u, v, c1).cost;
            } else {
              lo = c1;
              c1 = c2;
              f1 = f2;
              c2 = lo + inv_phi * (hi - lo);
              f2 = evaluate(&prob, u, v, c2).cost;
            }
          }
          const double cm = 0.5 * (lo + hi);
          const double fm = evaluate(&prob, u, v, cm).cost;
          if (fm < best_c) {
            best_c = fm;
            best_d = cm;
          }
          if (f1 < best_c) {
            best_c = f1;
            best_d = c1;
          }
          if (f2 < best_c) {
            best_c = f2;
            best_d = c2;
          }
        }
      }
      if (best_d < inv_min || best_d > inv_max) continue;
      const double fd = inv_to_depth(best_d);
      if (fd < cfg->min_depth_m || fd > cfg->max_depth_m) continue;
      depth[idx] = (float)fd;
      double pc = best_c;
      if (pc < 0.0) pc = 0.0;
      if (pc > 1.0) pc = 1.0;
      conf[idx] = (float)(1.0 - pc);
      ++refined;
    }
    free(fit_cost);
  }
  return refined;
}

/* ------------------------------------------------------------------------- */
/* FuseDepthMaps (patchmatch_stereo.cpp:985-1089)                            */
/* Output sorted by voxel key (the reference's order is unordered_map's).    */
/* ------------------------------------------------------------------------- */

typedef struct {
  int64_t key;
  int64_t order; /* (map, y, x) insertion order */
  double p[3];
  double n[3];
  float conf;
} fuse_entry;

static int cmp_entry(const void* a, const void* b) {
  const fuse_entry* x = (const fuse_entry*)a;
  const fuse_entry* y = (const fuse_entry*)b;
  if (x->key != y->key) return x->key < y->key ? -1 : 1;
  return x->order < y->order ? -1 : (x->order > y->order ? 1 : 0);
}

int64_t orc_fuse(int num_maps, const int32_t* widths, const int32_t* heights,
                 const float* const* depth, const uint8_t* const* support,
                 const float* const* conf, const recon_mvs_camera* cams,
                 const recon_mvs_config* cfg, int64_t capacity, double* out_pos,
                 double* out_nrm, float* out_conf, int64_t* out_keys) {
  const double voxel = cfg->fusion_voxel_size_m;
  if (num_maps <= 0 || voxel <= 0.0) return 0;
  size_t total = 0;
  for (int m = 0; m < num_maps; ++m) total += (size_t)widths[m] * heights[m];
  fuse_entry* e = (fuse_entry*)malloc((total ? total : 1) * sizeof(fuse_entry));
  size_t ne = 0;
  int64_t order = 0;
  for (int m = 0; m < num_maps; ++m) {
    if (!depth[m]) continue;
    const recon_mvs_camera* c = &cams[m];
    for (int y = 0; y < heights[m]; ++y) {
      for (int x = 0; x < widths[m]; ++x, ++order) {
        const size_t idx = (size_t)y * widths[m] + x;
        const float d = depth[m][idx];
        if (!isfinite(d)) continue;
        if (support[m][idx] < cfg->min_view_consistency) continue;
        if (cfg->photo_threshold < 1.0) {
          const float photo = 1.0f - conf[m][idx];
          if (photo > cfg->photo_threshold) continue;
        }
        double ray[3];
        orc_pixel_to_ray(c, x + 0.5, y + 0.5, ray);
        const double dd = (double)d;
        const double pc[3] = {ray[0] * dd, ray[1] * dd, ray[2] * dd};
        fuse_entry* q = &e[ne++];
        orc_cam_to_world(c, pc, q->p);
        double rt[3];
        orc_matTvec(c->R, ray, rt);
        const double nn = orc_norm3(rt);
        for (int j = 0; j < 3; ++j) q->n[j] = nn < 1e-15 ? 0.0 : rt[j] / nn;
        for (int j = 0; j < 3; ++j) q->n[j] = -q->n[j];
        const int64_t vx = (int64_t)floor(q->p[0] / voxel);
        const int64_t vy = (int64_t)floor(q->p[1] / voxel);
        const int64_t vz = (int64_t)floor(q->p[2] / voxel);
        q->key = (int64_t)(((uint64_t)vx * 73856093ULL) ^
                           ((uint64_t)vy * 19349663ULL) ^
                           ((uint64_t)vz * 83492791ULL));
        q->conf = conf[m][idx];
        q->order = order;
      }
    }
  }
  qsort(e, ne, sizeof(fuse_entry), cmp_entry);
  int64_t count = 0;
  size_t i = 0;
  while (i < ne) {
    size_t j = i;
    double sp[3] = {0.0, 0.0, 0.0}, sn[3] = {0.0, 0.0, 0.0}, sc = 0.0;
    uint32_t cnt = 0;
    while (j < ne && e[j].key == e[i].key) {
      for (int k = 0; k < 3; ++k) {
        sp[k] += e[j].p[k];
        sn[k] += e[j].n[k];
      }
      sc += e[j].conf;
      ++cnt;
      ++j;
    }
    if (count < capacity) {
      const double n = (double)cnt;
      double nrm[3] = {sn[0], sn[1], sn[2]};
      const double len = orc_norm3(nrm);
      if (len > 1e-12) {
        const double l2 = orc_norm3(nrm);
        if (!(l2 < 1e-15))
          for (int k = 0; k < 3; ++k) nrm[k] /= l2;
      } else {
        nrm[0] = 0.0;
        nrm[1] = 0.0;
        nrm[2] = 1.0;
      }
      double cf = (sc / n) * sqrt(n);
      if (cf > 1.0) cf = 1.0;
      for (int k = 0; k < 3; ++k) {
        out_pos[3 * count + k] = sp[k] / n;
        out_nrm[3 * count + k] = nrm[k];
      }
      out_conf[count] = (float)cf;
      if (out_keys) out_keys[count] = e[i].key;
    }
    ++count;
    i = j;
  }
  free(e);
  return count;
}

/* ------------------------------------------------------------------------- */
/* SelectSourceViews (patchmatch_stereo.cpp:133-211)                         */
/* ------------------------------------------------------------------------- */

typedef struct {
  int index;
  double score;
} view_cand;

int orc_select_views(const recon_mvs_camera* ref, int n, const recon_mvs_camera* poses,
                     const recon_mvs_config* cfg, int32_t* out, int capacity) {
  view_cand* c = (view_cand*)malloc((size_t)(n > 0 ? n : 1) * sizeof(view_cand));
  int nc = 0;
  double rc[3];
  orc_center(ref, rc);
  for (int i = 0; i < n; ++i) {
    if (i == cfg->ref_view_index) continue;
    double sc[3];
    orc_center(&poses[i], sc);
    const double b[3] = {sc[0] - rc[0], sc[1] - rc[1], sc[2] - rc[2]};
    if (orc_norm3(b) < 1e-6) continue;
    const double z_ref = cfg->reference_depth_m;
    const double z_src = z_ref + b[2];
    if (z_src <= 1e-6) continue;
    if (fabs(z_ref / z_src - 1.0) > cfg->max_view_scale_mismatch) continue;
    double rel[3];
    orc_matvec(ref->R, b, rel);
    const double dist = orc_norm3(rel);
    if (dist < 1e-6) continue;
    double cosv = rel[2] / dist;
    if (cosv < -1.0) cosv = -1.0;
    if (cosv > 1.0) cosv = 1.0;
    const double ang = acos(cosv) * 180.0 / 3.14159265358979323846;
    double score;
    if (ang < cfg->ideal_angle_min_deg) {
      score = ang / cfg->ideal_angle_min_deg;
    } else if (ang > cfg->ideal_angle_max_deg) {
      score = 1.0 - (ang - cfg->ideal_angle_max_deg) / (cfg->ideal_angle_max_deg * 2.0);
      if (score < 0.0) score = 0.0;
    } else {
      score = 1.0;
    }
    c[nc].index = i;
    c[nc].score = score * cfg->view_angle_weight;
    ++nc;
  }
  /* stable descending sort by score (insertion sort keeps input order on ties) */
  for (int i = 1; i < nc; ++i) {
    view_cand t = c[i];
    int j = i - 1;
    while (j >= 0 && c[j].score < t.score) {
      c[j + 1] = c[j];
      --j;
    }
    c[j + 1] = t;
  }
  int m = 0;
  for (int i = 0; i < nc && m < cfg->num_views; ++i) {
    if (m < capacity) out[m] = c[i].index;
    ++m;
  }
  free(c);
  return m;
}
