/*
 * pm_oracle_plane.c — TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * FP64 restatement of the plane-mode path of the GPU library (slanted planes,
 * homography warps, top-k NCC, red-black sweeps, geometric consistency).
 * The reference has no CPU code for these; they follow its design documents
 * (docs/05_单目管线.md §5.3.3-§5.6), and the special case "fronto-parallel
 * plane, translation-only cameras, k = all views, stride 1, reference
 * sampling convention" reduces exactly to EvaluateCandidate
 * (patchmatch_stereo.cpp:344-418), which tests/test_oracle_plane.py pins
 * against the compiled reference.
 *
 * Costs are evaluated in FP64 with the exact homography of every sample.
 * Candidate planes are generated with the same FP32 operations as the GPU
 * (this file is built with -ffp-contract=off), so both sides score the same
 * candidates; trajectories can then differ only where FP32 and FP64 costs
 * order two candidates differently.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "pm_oracle.h"

#define ORC_MAXV 16
#define ORC_COST_INF 1e30

/* lowbias32 (shaders/common/gpu_algo.h:150-154; shaders/metal/gpu_common.h:239-244). */
uint32_t orc_lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

static uint32_t rng_key(uint32_t seed, uint32_t pass, uint32_t pixel) {
  return orc_lowbias32(orc_lowbias32(orc_lowbias32(seed) ^ pass) ^ pixel);
}

static float rng_uniform(uint32_t key, uint32_t k) {
  return (float)(orc_lowbias32(key ^ ((k + 1u) * 0x9E3779B9u)) >> 8) * (1.0f / 16777216.0f);
}

float orc_rng_uniform(uint32_t seed, uint32_t pass, uint32_t pixel, uint32_t k) {
  return rng_uniform(rng_key(seed, pass, pixel), k);
}

typedef struct {
  int w, h;
  const float* ref;
  double fx, fy, cx, cy;
  float ffx, ffy, fcx, fcy;
  int nsrc, top_k;
  const float* src[ORC_MAXV];
  int sw[ORC_MAXV], sh[ORC_MAXV];
  double A[ORC_MAXV][9], b[ORC_MAXV][3];
  double Rrel[ORC_MAXV][9], trel[ORC_MAXV][3]; /* reference camera -> source camera */
  double sfx[ORC_MAXV], sfy[ORC_MAXV], scx[ORC_MAXV], scy[ORC_MAXV];
  float min_depth, max_depth, inv_min, inv_max;
  double eps, photo_reject;
  uint32_t seed;
  int reach;
  float depth_perturb, normal_perturb;
  int half, stride;
  double sample_shift; /* 0.5: pixel centres at integers (plane mode); 0: reference convention */
} plane_args;

typedef struct {
  float d, nx, ny, nz;
} hyp;

/* H = K_s R_rel K_r^-1, b = K_s t_rel with R_rel = R_s R_r^T, t_rel = t_s - R_rel t_r. */
static void view_pair(const recon_mvs_camera* R, const recon_mvs_camera* S, double A[9],
                      double b[3], double Rrel[9], double trel[3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l) s += S->R[3 * i + l] * R->R[3 * j + l];
      Rrel[3 * i + j] = s;
    }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int l = 0; l < 3; ++l) s += Rrel[3 * i + l] * R->t[l];
    trel[i] = S->t[i] - s;
  }
  const double Ks[9] = {S->fx, 0.0, S->cx, 0.0, S->fy, S->cy, 0.0, 0.0, 1.0};
  const double Ki[9] = {1.0 / R->fx, 0.0, -R->cx / R->fx, 0.0, 1.0 / R->fy, -R->cy / R->fy,
                        0.0, 0.0, 1.0};
  double KR[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l) s += Ks[3 * i + l] * Rrel[3 * l + j];
      KR[3 * i + j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l) s += KR[3 * i + l] * Ki[3 * l + j];
      A[3 * i + j] = s;
    }
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int l = 0; l < 3; ++l) s += Ks[3 * i + l] * trel[l];
    b[i] = s;
  }
}

/* Parameters: same layout as recon_mvs_plane_params in include/recon_mvs_hip.h. */
typedef struct {
  int32_t patch_size, window_stride, top_k, iterations;
  uint32_t seed;
  int32_t propagation_reach, refine_iterations, tile_schedule;
  double min_depth_m, max_depth_m, depth_perturb, normal_perturb, photo_threshold, value_scale;
} orc_plane_params;

static int make_args(plane_args* a, const float* const* imgs, const recon_mvs_camera* cams,
                     int ref, int nsrc, const int32_t* srcs, const orc_plane_params* p,
                     double sample_shift) {
  if (nsrc < 1 || nsrc > ORC_MAXV) return 0;
  memset(a, 0, sizeof(*a));
  const recon_mvs_camera* R = &cams[ref];
  a->w = R->width;
  a->h = R->height;
  a->ref = imgs[ref];
  a->fx = R->fx; a->fy = R->fy; a->cx = R->cx; a->cy = R->cy;
  a->ffx = (float)R->fx; a->ffy = (float)R->fy; a->fcx = (float)R->cx; a->fcy = (float)R->cy;
  a->nsrc = nsrc;
  a->top_k = p->top_k == 0 ? nsrc : (p->top_k < nsrc ? p->top_k : nsrc);
  for (int k = 0; k < nsrc; ++k) {
    const recon_mvs_camera* S = &cams[srcs[k]];
    a->src[k] = imgs[srcs[k]];
    a->sw[k] = S->width;
    a->sh[k] = S->height;
    view_pair(R, S, a->A[k], a->b[k], a->Rrel[k], a->trel[k]);
    a->sfx[k] = S->fx;
    a->sfy[k] = S->fy;
    a->scx[k] = S->cx;
    a->scy[k] = S->cy;
  }
  a->min_depth = (float)p->min_depth_m;
  a->max_depth = (float)p->max_depth_m;
  a->inv_min = (float)(1.0 / p->max_depth_m);
  a->inv_max = (float)(1.0 / p->min_depth_m);
  a->eps = 1e-9 * (p->value_scale / 255.0) * (p->value_scale / 255.0);
  a->photo_reject = 3.0 * p->photo_threshold;
  a->seed = p->seed;
  a->reach = p->propagation_reach;
  a->depth_perturb = (float)p->depth_perturb;
  a->normal_perturb = (float)p->normal_perturb;
  a->half = p->patch_size / 2;
  a->stride = p->window_stride;
  a->sample_shift = sample_shift;
  return 1;
}

/* Bilinear sample with the reference's semantics (patchmatch_stereo.cpp:78-100)
   on float pixels. */
static int sample_f(const float* img, int w, int h, double u, double v, double* out) {
  if (u < 0.0 || v < 0.0 || u > (double)(w - 1) || v > (double)(h - 1)) return 0;
  const int x0 = (int)floor(u), y0 = (int)floor(v);
  const int x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1;
  const int y1 = y0 + 1 < h - 1 ? y0 + 1 : h - 1;
  const double fx = u - x0, fy = v - y0;
  const double p00 = img[y0 * w + x0], p10 = img[y0 * w + x1];
  const double p01 = img[y1 * w + x0], p11 = img[y1 * w + x1];
  const double top = p00 * (1.0 - fx) + p10 * fx;
  const double bot = p01 * (1.0 - fx) + p11 * fx;
  *out = top * (1.0 - fy) + bot * fy;
  return 1;
}

/* FP64 plane cost; view_ncc (may be NULL) gets NaN for invalid views. */
static double plane_cost(const plane_args* a, int x, int y, hyp h, int* nvalid_out,
                         double* view_ncc) {
  *nvalid_out = 0;
  for (int v = 0; view_ncc && v < a->nsrc; ++v) view_ncc[v] = NAN;
  if (!(h.d >= a->min_depth && h.d <= a->max_depth)) return ORC_COST_INF;
  const double qx = x + 0.5, qy = y + 0.5;
  const double rx = (qx - a->cx) / a->fx, ry = (qy - a->cy) / a->fy;
  const double nx = h.nx, ny = h.ny, nz = h.nz, D = h.d;
  const double nd = nx * rx + ny * ry + nz;
  if (!(nd < 0.0)) return ORC_COST_INF;
  const double rho = D * nd;
  const double m0 = nx / a->fx / rho;
  const double m1 = ny / a->fy / rho;
  const double m2 = (nz - nx * a->cx / a->fx - ny * a->cy / a->fy) / rho;
  const int S1 = 2 * a->half / a->stride + 1;
  const int S = S1 * S1;
  double costs[ORC_MAXV];
  int nvalid = 0;
  for (int v = 0; v < a->nsrc; ++v) {
    double H[9];
    for (int i = 0; i < 3; ++i) {
      H[3 * i + 0] = a->A[v][3 * i + 0] + a->b[v][i] * m0;
      H[3 * i + 1] = a->A[v][3 * i + 1] + a->b[v][i] * m1;
      H[3 * i + 2] = a->A[v][3 * i + 2] + a->b[v][i] * m2;
    }
    const double hc2 = H[6] * qx + H[7] * qy + H[8];
    if (!(hc2 * D > 1e-6)) continue;
    const double hc0 = H[0] * qx + H[1] * qy + H[2], hc1 = H[3] * qx + H[4] * qy + H[5];
    if (!(fabs(hc0 / hc2 - 0.5) < 1e7 && fabs(hc1 / hc2 - 0.5) < 1e7)) continue;
    double sr = 0, ss = 0, srr = 0, sss = 0, srs = 0;
    int cnt = 0;
    /* two passes: means, then centred sums */
    double vr_[1024], vs_[1024];
    for (int j = 0; j < S1; ++j) {
      for (int i = 0; i < S1; ++i) {
        const int dx = -a->half + i * a->stride, dy = -a->half + j * a->stride;
        double rv;
        if (a->sample_shift == 0.5) {
          const int gx = x + dx, gy = y + dy;
          if (gx < 0 || gy < 0 || gx >= a->w || gy >= a->h) continue;
          rv = a->ref[gy * a->w + gx];
        } else {
          if (!sample_f(a->ref, a->w, a->h, x + 0.5 + dx - a->sample_shift,
                        y + 0.5 + dy - a->sample_shift, &rv))
            continue;
        }
        const double px = x + 0.5 + dx, py = y + 0.5 + dy;
        const double h2 = H[6] * px + H[7] * py + H[8];
        if (!(h2 > 0.0)) continue;
        const double u = (H[0] * px + H[1] * py + H[2]) / h2 - a->sample_shift;
        const double vv = (H[3] * px + H[4] * py + H[5]) / h2 - a->sample_shift;
        double sv;
        if (!sample_f(a->src[v], a->sw[v], a->sh[v], u, vv, &sv)) continue;
        vr_[cnt] = rv;
        vs_[cnt] = sv;
        sr += rv;
        ss += sv;
        ++cnt;
      }
    }
    if (cnt < S / 2 || cnt == 0) continue;
    const double mr = sr / cnt, ms = ss / cnt;
    for (int k = 0; k < cnt; ++k) {
      const double a1 = vr_[k] - mr, b1 = vs_[k] - ms;
      srr += a1 * a1;
      sss += b1 * b1;
      srs += a1 * b1;
    }
    const double var_r = srr / cnt, var_s = sss / cnt, cov = srs / cnt;
    const double denom = sqrt(var_r * var_s);
    if (!(denom >= a->eps) || !(denom > 0.0)) continue;
    double ncc = cov / denom;
    if (ncc < -1.0) ncc = -1.0;
    if (ncc > 1.0) ncc = 1.0;
    if (view_ncc) view_ncc[v] = ncc;
    costs[nvalid++] = 1.0 - ncc;
  }
  *nvalid_out = nvalid;
  if (nvalid == 0) return ORC_COST_INF;
  /* mean of the top_k smallest */
  for (int i = 1; i < nvalid; ++i) {
    const double t = costs[i];
    int j = i - 1;
    while (j >= 0 && costs[j] > t) {
      costs[j + 1] = costs[j];
      --j;
    }
    costs[j + 1] = t;
  }
  const int kk = a->top_k < nvalid ? a->top_k : nvalid;
  double sum = 0.0;
  for (int k = 0; k < kk; ++k) sum += costs[k];
  const double c = sum / kk;
  return c > a->photo_reject ? ORC_COST_INF : c;
}

/* ---- candidate planes, FP32, same operations as the GPU ---- */

static void pixel_ray(const plane_args* a, int x, int y, float r[3]) {
  r[0] = ((float)x + 0.5f - a->fcx) / a->ffx;
  r[1] = ((float)y + 0.5f - a->fcy) / a->ffy;
  r[2] = 1.0f;
}

static int normalize3(float n[3]) {
  const float l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (!(l2 > 1e-12f)) return 0;
  const float l = sqrtf(l2);
  n[0] = n[0] / l;
  n[1] = n[1] / l;
  n[2] = n[2] / l;
  return 1;
}

static hyp random_plane(const plane_args* a, uint32_t key, uint32_t k0, const float ray[3]) {
  const float u0 = rng_uniform(key, k0);
  const float inv = a->inv_min + u0 * (a->inv_max - a->inv_min);
  float n[3] = {2.0f * rng_uniform(key, k0 + 1) - 1.0f, 2.0f * rng_uniform(key, k0 + 2) - 1.0f,
                2.0f * rng_uniform(key, k0 + 3) - 1.0f};
  if (!normalize3(n)) {
    n[0] = 0.0f;
    n[1] = 0.0f;
    n[2] = -1.0f;
  }
  const float dp = n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2];
  if (dp > 0.0f) {
    n[0] = -n[0];
    n[1] = -n[1];
    n[2] = -n[2];
  }
  hyp h = {1.0f / inv, n[0], n[1], n[2]};
  return h;
}

static int transfer(const plane_args* a, const float p[4], int px, int py, const float ray[3],
                    hyp* out) {
  float rn[3];
  pixel_ray(a, px, py, rn);
  const float rho = p[0] * (p[1] * rn[0] + p[2] * rn[1] + p[3] * rn[2]);
  const float dn = p[1] * ray[0] + p[2] * ray[1] + p[3] * ray[2];
  if (!(dn < 0.0f)) return 0;
  const float d = rho / dn;
  if (!(d >= a->min_depth && d <= a->max_depth)) return 0;
  out->d = d;
  out->nx = p[1];
  out->ny = p[2];
  out->nz = p[3];
  return 1;
}

/* Exported: cost of n hypotheses (x, y, d, nx, ny, nz). */
int orc_plane_eval(int nviews, const float* const* imgs, const recon_mvs_camera* cams, int ref,
                   int nsrc, const int32_t* srcs, const orc_plane_params* p, double sample_shift,
                   int n, const float* hyps, double* cost, int32_t* nvalid, double* view_ncc) {
  (void)nviews;
  plane_args a;
  if (!make_args(&a, imgs, cams, ref, nsrc, srcs, p, sample_shift)) return 0;
#pragma omp parallel for schedule(dynamic, 16)
  for (int q = 0; q < n; ++q) {
    const float* hq = hyps + 6 * q;
    const int x = (int)hq[0], y = (int)hq[1];
    hyp h = {hq[2], hq[3], hq[4], hq[5]};
    int nv = 0;
    if (x < 0 || y < 0 || x >= a.w || y >= a.h) {
      cost[q] = ORC_COST_INF;
      nvalid[q] = 0;
      for (int v = 0; view_ncc && v < nsrc; ++v) view_ncc[q * nsrc + v] = NAN;
      continue;
    }
    cost[q] = plane_cost(&a, x, y, h, &nv, view_ncc ? view_ncc + (size_t)q * nsrc : NULL);
    nvalid[q] = nv;
  }
  return 1;
}

/* Exported: full plane-mode PatchMatch.  planes: P x 4 floats (d, nx, ny, nz);
   cost: P doubles; support: P bytes.  If init_planes is non-NULL the random
   init is replaced by those planes (then scored) and `iterations` sweeps run;
   rows whose depth is NaN keep the random plane (seeded init: only pixels
   with a seed are given). */
int orc_plane_patchmatch(int nviews, const float* const* imgs, const recon_mvs_camera* cams,
                         int ref, int nsrc, const int32_t* srcs, const orc_plane_params* p,
                         const float* init_planes, float* planes, double* cost,
                         uint8_t* support) {
  (void)nviews;
  plane_args a;
  if (!make_args(&a, imgs, cams, ref, nsrc, srcs, p, 0.5)) return 0;
  const int W = a.w, Hh = a.h;
  const int P = W * Hh;
#pragma omp parallel for schedule(dynamic, 64)
  for (int idx = 0; idx < P; ++idx) {
    const int x = idx % W, y = idx / W;
    float ray[3];
    pixel_ray(&a, x, y, ray);
    hyp h;
    if (init_planes && init_planes[4 * idx] == init_planes[4 * idx]) {
      h.d = init_planes[4 * idx];
      h.nx = init_planes[4 * idx + 1];
      h.ny = init_planes[4 * idx + 2];
      h.nz = init_planes[4 * idx + 3];
    } else {
      h = random_plane(&a, rng_key(a.seed, 0u, (uint32_t)idx), 0, ray);
    }
    int nv = 0;
    cost[idx] = plane_cost(&a, x, y, h, &nv, NULL);
    planes[4 * idx] = h.d;
    planes[4 * idx + 1] = h.nx;
    planes[4 * idx + 2] = h.ny;
    planes[4 * idx + 3] = h.nz;
    support[idx] = (uint8_t)(nv < 255 ? nv : 255);
  }
  for (int it = 0; it < p->iterations; ++it) {
    for (int color = 0; color < 2; ++color) {
      const float dscale = ldexpf(a.depth_perturb, -it);
      const float nscale = ldexpf(a.normal_perturb, -it);
#pragma omp parallel for schedule(dynamic, 64)
      for (int idx = 0; idx < P; ++idx) {
        const int x = idx % W, y = idx / W;
        if (((x + y) & 1) != color) continue;
        float ray[3];
        pixel_ray(&a, x, y, ray);
        float cur[4] = {planes[4 * idx], planes[4 * idx + 1], planes[4 * idx + 2], planes[4 * idx + 3]};
        const float base[4] = {cur[0], cur[1], cur[2], cur[3]};
        double ccur = cost[idx];
        int scur = support[idx];
        const uint32_t key = rng_key(a.seed, (uint32_t)(1 + 2 * it + color), (uint32_t)idx);
        int sx[4], sy[4];
        const int ddx[4] = {-1, 1, 0, 0}, ddy[4] = {0, 0, -1, 1};
        for (int d = 0; d < 4; ++d) {
          sx[d] = -1;
          sy[d] = -1;
          double best = ORC_COST_INF;
          for (int s = 0; s < a.reach; ++s) {
            const int dist = 2 * s + 1;
            const int nx = x + ddx[d] * dist, ny = y + ddy[d] * dist;
            if (nx < 0 || ny < 0 || nx >= W || ny >= Hh) break;
            const double c = cost[ny * W + nx];
            if (c < best) {
              best = c;
              sx[d] = nx;
              sy[d] = ny;
            }
          }
        }
        for (int c = 0; c < 8; ++c) {
          hyp h;
          int ok = 1;
          if (c < 4) {
            if (sx[c] < 0) continue;
            ok = transfer(&a, planes + 4 * (sy[c] * W + sx[c]), sx[c], sy[c], ray, &h);
          } else if (c < 7) {
            const uint32_t k0 = 4u * (uint32_t)(c - 4);
            float d = base[0];
            float n[3] = {base[1], base[2], base[3]};
            if (c != 5) {
              const float inv = 1.0f / d;
              const float u = 2.0f * rng_uniform(key, k0) - 1.0f;
              const float inv2 = inv + inv * (dscale * u);
              ok = inv2 >= a.inv_min && inv2 <= a.inv_max;
              d = 1.0f / inv2;
            }
            if (ok && c != 4) {
              n[0] = n[0] + nscale * (2.0f * rng_uniform(key, k0 + 1) - 1.0f);
              n[1] = n[1] + nscale * (2.0f * rng_uniform(key, k0 + 2) - 1.0f);
              n[2] = n[2] + nscale * (2.0f * rng_uniform(key, k0 + 3) - 1.0f);
              ok = normalize3(n) && (n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2]) < 0.0f;
            }
            h.d = d;
            h.nx = n[0];
            h.ny = n[1];
            h.nz = n[2];
          } else {
            h = random_plane(&a, key, 12u, ray);
          }
          if (!ok) continue;
          int nv = 0;
          const double cc = plane_cost(&a, x, y, h, &nv, NULL);
          if (cc < ccur) {
            ccur = cc;
            cur[0] = h.d;
            cur[1] = h.nx;
            cur[2] = h.ny;
            cur[3] = h.nz;
            scur = nv;
          }
        }
        for (int k = 0; k < 4; ++k) planes[4 * idx + k] = cur[k];
        cost[idx] = ccur;
        support[idx] = (uint8_t)(scur < 255 ? scur : 255);
      }
    }
  }
  return 1;
}

/* ---- cost and adjoint in point form (FP64) ----
   The sample at window offset (dx, dy) of pixel (x, y) lies on the ray
   r_i = K_r^-1 (x + 0.5 + dx, y + 0.5 + dy, 1) at X_i = lambda_i r_i, with
   lambda_i = D (n.r_c) / (n.r_i) (the plane through D r_c with normal n).  In
   the source camera Y_i = R_rel X_i + t_rel, and u = f_s Y_x / Y_z + c_s.  This
   is the map of plane_cost's homography, written independently of it.
   Source poses are perturbed on the left: Y -> Exp(w) Y + tau, so
   dY/dw = -[Y]x and dY/dtau = I.  Validity rules are those of plane_cost. */

/* Bilinear sample as sample_f, plus its derivatives in u and v (one-sided at
   the right and bottom edges, where x1 = x0 or y1 = y0). */
static int sample_fd(const float* img, int w, int h, double u, double v, double* out,
                     double* du, double* dv) {
  if (u < 0.0 || v < 0.0 || u > (double)(w - 1) || v > (double)(h - 1)) return 0;
  const int x0 = (int)floor(u), y0 = (int)floor(v);
  const int x1 = x0 + 1 < w - 1 ? x0 + 1 : w - 1;
  const int y1 = y0 + 1 < h - 1 ? y0 + 1 : h - 1;
  const double fx = u - x0, fy = v - y0;
  const double p00 = img[y0 * w + x0], p10 = img[y0 * w + x1];
  const double p01 = img[y1 * w + x0], p11 = img[y1 * w + x1];
  const double top = p00 * (1.0 - fx) + p10 * fx;
  const double bot = p01 * (1.0 - fx) + p11 * fx;
  *out = top * (1.0 - fy) + bot * fy;
  *du = (x1 > x0 ? 1.0 : 0.0) * ((p10 - p00) * (1.0 - fy) + (p11 - p01) * fy);
  *dv = (y1 > y0 ? 1.0 : 0.0) * (bot - top);
  return 1;
}

typedef struct {
  double rv, sv, du, dv;  /* reference value, source value, d(source)/du, d(source)/dv */
  double r[3], nr;        /* ray r_i and n . r_i */
  double hv[3];           /* Y_i / lambda_i */
} orc_sample;

/* Samples of view v for the plane (D, n) at pixel (x, y); returns the count,
   or -1 when the view is rejected before sampling. */
static int view_samples(const plane_args* a, int v, int x, int y, double D, const double n[3],
                        double rho, orc_sample* out) {
  const double* Rr = a->Rrel[v];
  const double* tr = a->trel[v];
  const double rc[3] = {(x + 0.5 - a->cx) / a->fx, (y + 0.5 - a->cy) / a->fy, 1.0};
  /* centre: Y_c / D */
  double hc[3];
  for (int i = 0; i < 3; ++i)
    hc[i] = Rr[3 * i] * rc[0] + Rr[3 * i + 1] * rc[1] + Rr[3 * i + 2] * rc[2] + tr[i] / D;
  if (!(hc[2] * D > 1e-6)) return -1;
  const double uc = a->sfx[v] * hc[0] / hc[2] + a->scx[v];
  const double vc = a->sfy[v] * hc[1] / hc[2] + a->scy[v];
  if (!(fabs(uc - 0.5) < 1e7 && fabs(vc - 0.5) < 1e7)) return -1;
  const int S1 = 2 * a->half / a->stride + 1;
  int cnt = 0;
  for (int j = 0; j < S1; ++j) {
    for (int i = 0; i < S1; ++i) {
      const int dx = -a->half + i * a->stride, dy = -a->half + j * a->stride;
      const int gx = x + dx, gy = y + dy;
      if (gx < 0 || gy < 0 || gx >= a->w || gy >= a->h) continue;
      orc_sample* sm = &out[cnt];
      sm->r[0] = (x + 0.5 + dx - a->cx) / a->fx;
      sm->r[1] = (y + 0.5 + dy - a->cy) / a->fy;
      sm->r[2] = 1.0;
      sm->nr = n[0] * sm->r[0] + n[1] * sm->r[1] + n[2] * sm->r[2];
      const double q = sm->nr / rho;
      for (int k = 0; k < 3; ++k)
        sm->hv[k] = Rr[3 * k] * sm->r[0] + Rr[3 * k + 1] * sm->r[1] + Rr[3 * k + 2] * sm->r[2] + tr[k] * q;
      if (!(sm->hv[2] > 0.0)) continue;
      const double u = a->sfx[v] * sm->hv[0] / sm->hv[2] + a->scx[v];
      const double vv = a->sfy[v] * sm->hv[1] / sm->hv[2] + a->scy[v];
      if (!sample_fd(a->src[v], a->sw[v], a->sh[v], u - 0.5, vv - 0.5, &sm->sv, &sm->du, &sm->dv))
        continue;
      sm->rv = a->ref[gy * a->w + gx];
      ++cnt;
    }
  }
  return cnt;
}

typedef struct {
  double ncc, raw, mr, ms, sr, ss; /* clamped and raw NCC, means, standard deviations */
  int n;
} orc_stats;

static int view_stats(const plane_args* a, const orc_sample* sm, int cnt, orc_stats* st) {
  const int S1 = 2 * a->half / a->stride + 1;
  const int S = S1 * S1;
  if (cnt < S / 2 || cnt == 0) return 0;
  double sr = 0, ss = 0;
  for (int k = 0; k < cnt; ++k) {
    sr += sm[k].rv;
    ss += sm[k].sv;
  }
  const double mr = sr / cnt, ms = ss / cnt;
  double srr = 0, sss = 0, srs = 0;
  for (int k = 0; k < cnt; ++k) {
    const double a1 = sm[k].rv - mr, b1 = sm[k].sv - ms;
    srr += a1 * a1;
    sss += b1 * b1;
    srs += a1 * b1;
  }
  const double var_r = srr / cnt, var_s = sss / cnt, cov = srs / cnt;
  const double denom = sqrt(var_r * var_s);
  if (!(denom >= a->eps) || !(denom > 0.0)) return 0;
  st->raw = cov / denom;
  st->ncc = st->raw < -1.0 ? -1.0 : (st->raw > 1.0 ? 1.0 : st->raw);
  st->mr = mr;
  st->ms = ms;
  st->sr = sqrt(var_r);
  st->ss = sqrt(var_s);
  st->n = cnt;
  return 1;
}

#define ORC_MAX_SAMPLES 1024

/* Cost of plane (D, n) at (x, y) and, when g_plane is non-NULL, the gradient
   of wgt * cost w.r.t. (D, n) (into g_plane) and the source poses (added to
   g_pose[v][0..5] = (w, tau)).  Invalid or rejected: ORC_COST_INF, zero
   gradient. */
static double point_cost_grad(const plane_args* a, int x, int y, double D, const double n[3],
                              double wgt, int* nvalid_out, double g_plane[4], double (*g_pose)[6],
                              orc_sample* sm) {
  *nvalid_out = 0;
  if (g_plane) g_plane[0] = g_plane[1] = g_plane[2] = g_plane[3] = 0.0;
  if (!(D >= (double)a->min_depth && D <= (double)a->max_depth)) return ORC_COST_INF;
  const double rc[3] = {(x + 0.5 - a->cx) / a->fx, (y + 0.5 - a->cy) / a->fy, 1.0};
  const double nd = n[0] * rc[0] + n[1] * rc[1] + n[2] * rc[2];
  if (!(nd < 0.0)) return ORC_COST_INF;
  const double rho = D * nd;
  double cst[ORC_MAXV];
  int vid[ORC_MAXV];
  int nvalid = 0;
  for (int v = 0; v < a->nsrc; ++v) {
    const int cnt = view_samples(a, v, x, y, D, n, rho, sm);
    orc_stats st;
    if (cnt < 0 || !view_stats(a, sm, cnt, &st)) continue;
    cst[nvalid] = 1.0 - st.ncc;
    vid[nvalid] = v;
    ++nvalid;
  }
  *nvalid_out = nvalid;
  if (nvalid == 0) return ORC_COST_INF;
  /* stable sort of (cost, view) by cost */
  for (int i = 1; i < nvalid; ++i) {
    const double t = cst[i];
    const int tv = vid[i];
    int j = i - 1;
    while (j >= 0 && cst[j] > t) {
      cst[j + 1] = cst[j];
      vid[j + 1] = vid[j];
      --j;
    }
    cst[j + 1] = t;
    vid[j + 1] = tv;
  }
  const int kk = a->top_k < nvalid ? a->top_k : nvalid;
  double sum = 0.0;
  for (int k = 0; k < kk; ++k) sum += cst[k];
  const double c = sum / kk;
  if (c > a->photo_reject) return ORC_COST_INF;
  if (!g_plane) return c;
  double gD = 0.0, gn[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < kk; ++k) {
    const int v = vid[k];
    const int cnt = view_samples(a, v, x, y, D, n, rho, sm);
    orc_stats st;
    if (cnt < 0 || !view_stats(a, sm, cnt, &st)) continue;
    if (st.raw < -1.0 || st.raw > 1.0) continue; /* clamped: flat */
    const double dc = -wgt / kk;                  /* d(wgt cost) / d ncc_v */
    const double al = 1.0 / (st.n * st.sr * st.ss), be = st.ncc / (st.n * st.ss * st.ss);
    const double* Rr = a->Rrel[v];
    for (int q = 0; q < cnt; ++q) {
      const orc_sample* s = &sm[q];
      /* d ncc / d s_i = (r~_i / (sr ss) - ncc s~_i / ss^2) / N */
      const double gs = dc * ((s->rv - st.mr) * al - (s->sv - st.ms) * be);
      const double gu = gs * s->du, gv = gs * s->dv;
      const double lam = rho / s->nr;
      const double Y[3] = {lam * s->hv[0], lam * s->hv[1], lam * s->hv[2]};
      const double gY[3] = {gu * a->sfx[v] / Y[2], gv * a->sfy[v] / Y[2],
                            -(gu * a->sfx[v] * Y[0] + gv * a->sfy[v] * Y[1]) / (Y[2] * Y[2])};
      if (g_pose) {
        g_pose[v][0] += Y[1] * gY[2] - Y[2] * gY[1];
        g_pose[v][1] += Y[2] * gY[0] - Y[0] * gY[2];
        g_pose[v][2] += Y[0] * gY[1] - Y[1] * gY[0];
        g_pose[v][3] += gY[0];
        g_pose[v][4] += gY[1];
        g_pose[v][5] += gY[2];
      }
      double gX[3];
      for (int i = 0; i < 3; ++i) gX[i] = Rr[i] * gY[0] + Rr[3 + i] * gY[1] + Rr[6 + i] * gY[2];
      const double ai = gX[0] * s->r[0] + gX[1] * s->r[1] + gX[2] * s->r[2];
      gD += ai * lam / D;
      const double f = ai / s->nr;
      for (int i = 0; i < 3; ++i) gn[i] += f * (D * rc[i] - lam * s->r[i]);
    }
  }
  g_plane[0] = gD;
  g_plane[1] = gn[0];
  g_plane[2] = gn[1];
  g_plane[3] = gn[2];
  return c;
}

/* Exported: costs of n planes (FP64 inputs) at pixels xy, in point form; with
   grad_plane non-NULL also the gradients of sum_q weights[q] * cost[q] w.r.t.
   each plane (n x 4: D, n) and, with grad_pose non-NULL, w.r.t. the source
   poses (nsrc x 6: rotation w, translation tau; left perturbation).  Invalid
   planes: cost ORC_COST_INF, zero gradient.  The pose sum is reduced in a
   fixed order (chunks of 64 planes, chunk partials in chunk order). */
int orc_plane_cost_grad(int nviews, const float* const* imgs, const recon_mvs_camera* cams,
                        int ref, int nsrc, const int32_t* srcs, const orc_plane_params* p, int n,
                        const int32_t* xy, const double* planes, const double* weights,
                        double* cost, int32_t* nvalid, double* grad_plane, double* grad_pose) {
  (void)nviews;
  plane_args a;
  if (!make_args(&a, imgs, cams, ref, nsrc, srcs, p, 0.5)) return 0;
  const int S1 = 2 * a.half / a.stride + 1;
  if (S1 * S1 > ORC_MAX_SAMPLES) return 0;
  const int chunk = 64;
  const int nchunks = (n + chunk - 1) / chunk;
  double(*part)[ORC_MAXV][6] = NULL;
  if (grad_pose) {
    part = calloc((size_t)(nchunks > 0 ? nchunks : 1), sizeof(*part));
    if (!part) return 0;
  }
#pragma omp parallel
  {
    orc_sample* sm = malloc(sizeof(orc_sample) * ORC_MAX_SAMPLES);
#pragma omp for schedule(dynamic, 1)
    for (int ck = 0; ck < nchunks; ++ck) {
      for (int q = ck * chunk; q < n && q < (ck + 1) * chunk; ++q) {
        const int x = xy[2 * q], y = xy[2 * q + 1];
        int nv = 0;
        double g[4] = {0, 0, 0, 0};
        if (x < 0 || y < 0 || x >= a.w || y >= a.h || !sm) {
          cost[q] = ORC_COST_INF;
        } else {
          const double nn[3] = {planes[4 * q + 1], planes[4 * q + 2], planes[4 * q + 3]};
          const double w = weights ? weights[q] : 1.0;
          cost[q] = point_cost_grad(&a, x, y, planes[4 * q], nn, w, &nv, grad_plane ? g : NULL,
                                    part ? part[ck] : NULL, sm);
        }
        if (nvalid) nvalid[q] = nv;
        if (grad_plane)
          for (int k = 0; k < 4; ++k) grad_plane[4 * q + k] = g[k];
      }
    }
    free(sm);
  }
  if (grad_pose) {
    for (int v = 0; v < nsrc; ++v)
      for (int k = 0; k < 6; ++k) {
        double s = 0.0;
        for (int ck = 0; ck < nchunks; ++ck) s += part[ck][v][k];
        grad_pose[6 * v + k] = s;
      }
    free(part);
  }
  return 1;
}

/* Exported: geometric consistency (docs/05 §5.6.2), FP64, reference camera
   arithmetic.  depth maps are float (NaN = invalid); geo_support gets
   1 + number of other maps agreeing within `thr` (0 for invalid pixels). */
void orc_geo_consistency(int nmaps, const float* const* depth, const recon_mvs_camera* cams,
                         double thr, uint8_t* const* geo_support) {
  for (int m = 0; m < nmaps; ++m) {
    const recon_mvs_camera* me = &cams[m];
    const int W = me->width, Hh = me->height;
#pragma omp parallel for schedule(dynamic, 8)
    for (int y = 0; y < Hh; ++y) {
      for (int x = 0; x < W; ++x) {
        const int idx = y * W + x;
        const float d = depth[m][idx];
        int support = 0;
        if (isfinite(d)) {
          double ray[3];
          orc_pixel_to_ray(me, x + 0.5, y + 0.5, ray);
          const double dd = (double)d;
          const double pc[3] = {ray[0] * dd, ray[1] * dd, ray[2] * dd};
          double pw[3];
          orc_cam_to_world(me, pc, pw);
          support = 1;
          for (int o = 0; o < nmaps; ++o) {
            if (o == m) continue;
            const recon_mvs_camera* oc = &cams[o];
            double ps[3];
            orc_world_to_cam(oc, pw, ps);
            if (ps[2] <= 1e-6) continue;
            double u, v;
            if (!orc_project(oc, ps, &u, &v)) continue;
            const double fu = floor(u), fv = floor(v);
            if (fu < 0.0 || fv < 0.0 || fu >= oc->width || fv >= oc->height) continue;
            const float zq = depth[o][(int)fv * oc->width + (int)fu];
            if (!isfinite(zq)) continue;
            const double zm = (double)zq;
            const double e = fabs(ps[2] - zm) / fmax(ps[2], zm);
            if (e <= thr) ++support;
          }
        }
        geo_support[m][idx] = (uint8_t)(support < 255 ? support : 255);
      }
    }
  }
}

/* ---- sub-pixel refinement along the ray (plane mode) ----
   The search of the GPU's k_plane_refine for one pixel with its normal held:
   a 5-point least-squares parabola in inverse depth (samples inv0 + (k - 2)
   step, step = inv_per_px / 2, inv_per_px = 1 / (f B), the reference's step,
   patchmatch_stereo.cpp:840-841), whose vertex -- scaled by the step, where
   the reference adds it in index units (line 898) -- is evaluated when it
   lies within the samples; then golden section over best +- inv_per_px
   (docs/05_单目管线.md:1694-1700), and the best of the costs seen.
   Candidate inverse depths are formed with the GPU's FP32 operations; costs
   and the parabola fit are FP64.  The search is written over a cost callback
   so that tests can run it on closed-form functions. */
typedef double (*refine_cost_fn)(const void* ctx, float inv, int* nv);

static void refine_1d(refine_cost_fn f, const void* fc, float inv0, double c0, int s0,
                      float inv_per_px, int iterations, float inv_min, float inv_max,
                      float* best_inv, double* best_cost, int* best_nv) {
  float bi = inv0;
  double bc = c0;
  int bs = s0;
  const float step = 0.5f * inv_per_px;
  double fv[5];
  int fit_ok = 1;
  for (int k = 0; k < 5; ++k) {
    if (k == 2) {
      fv[k] = c0;
      continue;
    }
    int nv = 0;
    fv[k] = f(fc, inv0 + (float)(k - 2) * step, &nv);
    fit_ok = fit_ok && fv[k] < ORC_COST_INF;
  }
  if (fit_ok) {
    /* normal equations of c(x) = a x^2 + b x + c at x = -2..2 */
    const double aa = (2.0 * fv[0] - fv[1] - 2.0 * fv[2] - fv[3] + 2.0 * fv[4]) / 14.0;
    const double bb = (-2.0 * fv[0] - fv[1] + fv[3] + 2.0 * fv[4]) / 10.0;
    if (aa > 1e-12) {
      const float vtx = (float)(-bb / (2.0 * aa));
      if (vtx >= -2.0f && vtx <= 2.0f) {
        const float cand = inv0 + vtx * step;
        int nv = 0;
        const double cc = f(fc, cand, &nv);
        if (cc < bc) {
          bc = cc;
          bi = cand;
          bs = nv;
        }
      }
    }
  }
  float lo = fmaxf(bi - inv_per_px, inv_min), hi = fminf(bi + inv_per_px, inv_max);
  if (hi > lo && iterations > 0) {
    const float ip = 0.6180339887498949f;
    float c1 = hi - ip * (hi - lo), c2 = lo + ip * (hi - lo);
    int n1 = 0, n2 = 0;
    double f1 = f(fc, c1, &n1), f2 = f(fc, c2, &n2);
    for (int k = 0; k < iterations; ++k) {
      if (f1 < f2) {
        hi = c2;
        c2 = c1;
        f2 = f1;
        n2 = n1;
        c1 = hi - ip * (hi - lo);
        f1 = f(fc, c1, &n1);
      } else {
        lo = c1;
        c1 = c2;
        f1 = f2;
        n1 = n2;
        c2 = lo + ip * (hi - lo);
        f2 = f(fc, c2, &n2);
      }
    }
    if (f1 < bc) {
      bc = f1;
      bi = c1;
      bs = n1;
    }
    if (f2 < bc) {
      bc = f2;
      bi = c2;
      bs = n2;
    }
  }
  *best_inv = bi;
  *best_cost = bc;
  *best_nv = bs;
}

typedef struct {
  const plane_args* a;
  int x, y;
  float nx, ny, nz;
} pixel_refine_ctx;

static double pixel_refine_cost(const void* ctx, float inv, int* nv) {
  const pixel_refine_ctx* c = (const pixel_refine_ctx*)ctx;
  *nv = 0;
  if (!(inv >= c->a->inv_min && inv <= c->a->inv_max)) return ORC_COST_INF;
  hyp h = {1.0f / inv, c->nx, c->ny, c->nz};
  return plane_cost(c->a, c->x, c->y, h, nv, NULL);
}

/* Exported: plane-mode sub-pixel refinement of a depth map in place.
   planes P x 4 (d, nx, ny, nz), cost P (the cost of each plane; >= 1e29:
   invalid, left alone), support P.  inv_per_px = 1 / (f B) in FP32 with f
   the mean reference focal length and B the mean reference-source baseline
   (patchmatch_stereo.cpp:472-479, 840). */
int orc_plane_refine(int nviews, const float* const* imgs, const recon_mvs_camera* cams, int ref,
                     int nsrc, const int32_t* srcs, const orc_plane_params* p, float* planes,
                     double* cost, uint8_t* support) {
  (void)nviews;
  plane_args a;
  if (!make_args(&a, imgs, cams, ref, nsrc, srcs, p, 0.5)) return 0;
  double rc[3];
  orc_center(&cams[ref], rc);
  double bsum = 0.0;
  for (int k = 0; k < nsrc; ++k) {
    double sc[3];
    orc_center(&cams[srcs[k]], sc);
    const double d[3] = {sc[0] - rc[0], sc[1] - rc[1], sc[2] - rc[2]};
    bsum += orc_norm3(d);
  }
  const double baseline = bsum / (double)nsrc;
  if (!(baseline > 1e-9)) return 0;
  const double focal = 0.5 * (cams[ref].fx + cams[ref].fy);
  const float inv_per_px = (float)(1.0 / (focal * baseline));
  const int P = a.w * a.h;
#pragma omp parallel for schedule(dynamic, 64)
  for (int idx = 0; idx < P; ++idx) {
    if (!(cost[idx] < ORC_COST_INF)) continue;
    float* pl = planes + 4 * idx;
    pixel_refine_ctx c = {&a, idx % a.w, idx / a.w, pl[1], pl[2], pl[3]};
    float bi;
    double bc;
    int bs;
    refine_1d(pixel_refine_cost, &c, 1.0f / pl[0], cost[idx], support[idx], inv_per_px,
              p->refine_iterations, a.inv_min, a.inv_max, &bi, &bc, &bs);
    pl[0] = 1.0f / bi;
    cost[idx] = bc;
    support[idx] = (uint8_t)(bs < 255 ? bs : 255);
  }
  return 1;
}

/* Closed-form costs for pinning refine_1d: 0 = A (x - m)^2 (a parabola: the
   fit recovers its vertex), 1 = A |x - m| (unimodal, not smooth: only the
   golden section finds it), 2 = 1 - exp(-((x - m) / A)^2) (unimodal, smooth,
   not quadratic); infinite outside [lo, hi], as a pixel's cost is outside the
   inverse-depth range. */
typedef struct {
  int kind;
  double m, A;
  float lo, hi;
  int evals;
} closed_form_ctx;

static double closed_form_cost(const void* ctx, float inv, int* nv) {
  closed_form_ctx* c = (closed_form_ctx*)ctx;
  ++c->evals;
  *nv = 0;
  if (!(inv >= c->lo && inv <= c->hi)) return ORC_COST_INF;
  *nv = 1;
  const double x = (double)inv - c->m;
  if (c->kind == 0) return c->A * x * x;
  if (c->kind == 1) return c->A * fabs(x);
  return 1.0 - exp(-(x / c->A) * (x / c->A));
}

/* Exported: refine_1d on a closed-form cost; returns the number of cost
   evaluations. */
int orc_refine_1d_closed_form(int kind, double m, double A, float inv0, float inv_per_px,
                              int iterations, float inv_min, float inv_max, float* best_inv,
                              double* best_cost) {
  closed_form_ctx c = {kind, m, A, inv_min, inv_max, 0};
  int nv0 = 0, bs = 0;
  const double c0 = closed_form_cost(&c, inv0, &nv0);
  c.evals = 0;
  refine_1d(closed_form_cost, &c, inv0, c0, nv0, inv_per_px, iterations, inv_min, inv_max,
            best_inv, best_cost, &bs);
  return c.evals;
}
