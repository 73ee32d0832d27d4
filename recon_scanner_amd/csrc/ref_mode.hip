// ref_mode.hip — reference mode: the reference CPU PatchMatch reproduced on
// the GPU in FP64 (compiled with -ffp-contract=off, so every double operation
// rounds exactly as in the reference's plain x86-64 build).
//
//   ComputeNcc        patchmatch_stereo.cpp:214-274  -> NccU8 / k_compute_ncc
//   EvaluateCandidate patchmatch_stereo.cpp:344-418  -> Evaluate / k_evaluate
//   PatchMatchStereo  patchmatch_stereo.cpp:422-764  -> k_pm_init_cost, k_pm_pass,
//                                                      k_pm_restart, k_pm_output
//                                                      + host driver
//   RefineSubpixel    patchmatch_stereo.cpp:770-979  -> k_refine
//
// The reference's directional passes read neighbours only from a snapshot of
// the previous pass (lines 636-685), so every pixel of a pass is independent:
// one GPU thread per pixel, ping-pong state buffers.  Its random numbers come
// from sequential xorshift64* streams, one per row block and iteration
// (lines 647-649), consumed in the pass's scan order; the host replays those
// streams into a per-pass table (draws are data-independent), so the GPU
// result equals the reference's for the same config.num_threads.
//
// The pixel kernels (initial costs, passes, refinement) change only where the
// values come from, never an operation: a workgroup owns a 32x4 tile whose
// reference patch samples are computed once into LDS; sources are read as
// 4-tap quads (one load per sample); a wave whose windows all lie inside the
// images takes a path without per-sample tests and with the reference sums
// precomputed per pixel, the others a masked path in which an invalid pair
// adds +0.0 to every sum.
#include <algorithm>
#include <cstring>
#include <thread>
#include <type_traits>

#include "mvs_common.hpp"

namespace rmvs {

constexpr int kMaxSrc = 32;
constexpr int kMaxFitPoints = 31;
// Timer classes (recon_mvs_hip_timing) of the reference-mode kernels: a
// directional pass counts as a sweep, the initial costs as init.
constexpr int kTimerPass = 0;
constexpr int kTimerInit = 1;
constexpr int kTimerRefine = 2;

// A source texel of the quad images (see MakeQuad).
using Quad = uint32_t;
constexpr int kQuadBytes = sizeof(Quad);

struct RefProblem {
  const uint8_t* ref;
  int rw, rh;
  int nsrc;
  int patch;
  const uint8_t* src[kMaxSrc];
  const Quad* srcq[kMaxSrc];  // source k as quads (k_make_quads)
  int sw[kMaxSrc];
  int sh[kMaxSrc];
  const recon_mvs_camera* cams;  // [0] reference, [1 + k] source k
  double min_depth, max_depth, photo_threshold;
};

// Pixel kernels run on TW x TH tiles: the reference patch samples of a tile
// (a HALF-pixel halo around it) are computed once into LDS.
constexpr int kTileW = 32;
constexpr int kTileH = 4;
constexpr int kTileThreads = kTileW * kTileH;
// tiles per block of the blocked tile order (LoadRefTile): 64 x 128 pixels
// (measured at 1280x960 against 8 sources: 262 ms per view; whole tile rows
// 283, 256 x 64 268, 128 x 128 265, 64 x 256 262)
constexpr int kBlockW = 2;
constexpr int kBlockH = 32;

// SampleBilinear (patchmatch_stereo.cpp:78-100): integer coordinates are pixel
// centres, out-of-range samples are rejected, not clamped.
__device__ __forceinline__ bool SampleU8(const uint8_t* __restrict__ img, int w, int h, double u,
                                         double v, double* out) {
  if (u < 0.0 || v < 0.0 || u > static_cast<double>(w - 1) || v > static_cast<double>(h - 1))
    return false;
  const int x0 = static_cast<int>(floor(u));
  const int y0 = static_cast<int>(floor(v));
  const int x1 = min(x0 + 1, w - 1);
  const int y1 = min(y0 + 1, h - 1);
  const double fx = u - x0;
  const double fy = v - y0;
  const double p00 = img[y0 * w + x0];
  const double p10 = img[y0 * w + x1];
  const double p01 = img[y1 * w + x0];
  const double p11 = img[y1 * w + x1];
  const double top = p00 * (1.0 - fx) + p10 * fx;
  const double bot = p01 * (1.0 - fx) + p11 * fx;
  *out = top * (1.0 - fy) + bot * fy;
  return true;
}

// Quad image of a source: texel (y, x) holds the four pixels a sample with
// floor (x, y) reads — p00, p10, p01, p11, the clamps of x1 and y1 applied
// when the quads were built (k_make_quads) — as bytes p00 | p10 << 8 |
// p01 << 16 | p11 << 24: one 4-byte load per sample instead of four 1-byte
// loads.
__device__ __forceinline__ Quad MakeQuad(uint8_t p00, uint8_t p10, uint8_t p01, uint8_t p11) {
  return static_cast<uint32_t>(p00) | static_cast<uint32_t>(p10) << 8 |
         static_cast<uint32_t>(p01) << 16 | static_cast<uint32_t>(p11) << 24;
}

// Tap i (0 p00, 1 p10, 2 p01, 3 p11) as a double.
__device__ __forceinline__ double QuadTap(Quad t, int i) {
  return static_cast<double>((t >> (8 * i)) & 0xffu);
}

// The texel at byte offset `off` through a raw buffer descriptor.
__device__ __forceinline__ Quad LoadQuad(__amdgpu_buffer_rsrc_t rsrc, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0);
}

// SampleQuad's arithmetic for a sample inside the source (no bounds tests).
__device__ __forceinline__ double Bilinear(Quad t, double fx, double fy) {
  const double top = QuadTap(t, 0) * (1.0 - fx) + QuadTap(t, 1) * fx;
  const double bot = QuadTap(t, 2) * (1.0 - fx) + QuadTap(t, 3) * fx;
  return top * (1.0 - fy) + bot * fy;
}

// SampleU8 on a quad image: same operations on the same values.
__device__ __forceinline__ bool SampleQuad(const Quad* __restrict__ q, int w, int h, double u,
                                           double v, double* out) {
  if (u < 0.0 || v < 0.0 || u > static_cast<double>(w - 1) || v > static_cast<double>(h - 1))
    return false;
  const int x0 = static_cast<int>(floor(u));
  const int y0 = static_cast<int>(floor(v));
  *out = Bilinear(q[y0 * w + x0], u - x0, v - y0);
  return true;
}

struct Ncc {
  bool valid;
  bool weak;
  int valid_samples;
  double ncc;
};

// Statistics of the sample pairs, in the reference's scan order (dy outer,
// dx inner) and accumulation order (patchmatch_stereo.cpp:231-250).
struct NccSums {
  double s_r = 0.0, s_s = 0.0, s_r2 = 0.0, s_s2 = 0.0, s_x = 0.0;
  int valid = 0;
  __device__ __forceinline__ void add(double vr, double vs) {
    s_r += vr;
    s_s += vs;
    s_r2 += vr * vr;
    s_s2 += vs * vs;
    s_x += vr * vs;
    ++valid;
  }
};

__device__ Ncc FinishNcc(const NccSums& a, int patch);

__device__ Ncc NccU8(const uint8_t* __restrict__ a, int aw, int ah, const uint8_t* __restrict__ b,
                     int bw, int bh, double ru, double rv, double su, double sv, int patch) {
  const int half = patch / 2;
  NccSums acc;
  for (int dy = -half; dy <= half; ++dy) {
    for (int dx = -half; dx <= half; ++dx) {
      double vr = 0.0, vs = 0.0;
      if (!SampleU8(a, aw, ah, ru + dx, rv + dy, &vr)) continue;
      if (!SampleU8(b, bw, bh, su + dx, sv + dy, &vs)) continue;
      acc.add(vr, vs);
    }
  }
  return FinishNcc(acc, patch);
}

// NccU8 for a reference patch centred on a pixel centre, its samples read
// from the LDS tile of RefTile (row `trow` of the tile, column `tcol`
// holding dx = -half; NaN = outside the reference image) and the source's
// from its quads.  A sample pair counts when both samples are inside, as in
// NccU8; the sampling has no side effects, so the order of the two tests does
// not matter.
__device__ Ncc NccTile(const double* __restrict__ tile, int tw, int trow, int tcol,
                       const Quad* __restrict__ q, int bw, int bh, double su, double sv,
                       int patch) {
  const int half = patch / 2;
  NccSums acc;
  for (int dy = -half; dy <= half; ++dy) {
    const double* row = tile + (trow + dy + half) * tw + tcol;
    for (int dx = -half; dx <= half; ++dx) {
      const double vr = row[dx + half];
      if (vr != vr) continue;
      double vs = 0.0;
      if (!SampleQuad(q, bw, bh, su + dx, sv + dy, &vs)) continue;
      acc.add(vr, vs);
    }
  }
  return FinishNcc(acc, patch);
}

__device__ Ncc FinishNcc(const NccSums& a, int patch) {
  Ncc r{false, false, 0, 0.0};
  const int valid = a.valid;
  const double s_r = a.s_r, s_s = a.s_s, s_r2 = a.s_r2, s_s2 = a.s_s2, s_x = a.s_x;
  if (valid < patch * patch / 2) return r;
  const double n = static_cast<double>(valid);
  const double mr = s_r / n;
  const double ms = s_s / n;
  const double var_r = s_r2 / n - mr * mr;
  const double var_s = s_s2 / n - ms * ms;
  r.weak = var_r < 12.0;  // kDefaultWeakTextureVariance (line 38)
  const double denom = sqrt(var_r * var_s);
  if (denom < 1e-9) return r;
  double ncc = (s_x / n - mr * ms) / denom;
  ncc = fmin(fmax(ncc, -1.0), 1.0);
  r.ncc = ncc;
  r.valid = true;
  r.valid_samples = valid;
  return r;
}

struct Cost {
  double cost, ncc_sum, weight_sum;
  int weak, valid_views;
};

// EvaluateCandidate with the reference ray precomputed by the caller; `ncc(k,
// su, sv)` is the NCC of the reference patch against source k at (su, sv).
//
// reject_above: a caller that only asks whether the cost is below some c
// passes c + kRejectMargin, and gets cost 1e30 as soon as the views left
// cannot bring the cost down to it.  Each view's NCC is at most 1, so with
// S the NCC sum over the n valid views so far and R views left the final
// cost is at least 1 - (S + R) / (n + R) (that bound grows with the number
// of valid views among the R); the margin, many orders above the rounding
// of these sums, keeps the test on the safe side.
constexpr double kRejectMargin = 1e-9;

template <class NccFn>
__device__ __forceinline__ Cost EvaluateWith(const RefProblem& p, const double* ray, double inv,
                                             NccFn ncc, double reject_above = 1e300) {
  Cost r{1e30, 0.0, 0.0, 0, 0};
  if (inv <= 1e-9) return r;
  const double depth = 1.0 / inv;
  if (depth < p.min_depth || depth > p.max_depth) return r;
  const double pc[3] = {ray[0] * depth, ray[1] * depth, ray[2] * depth};
  double pw[3];
  CamToWorld(p.cams[0], pc, pw);
  for (int k = 0; k < p.nsrc; ++k) {
    const double rem = static_cast<double>(p.nsrc - k);
    if (k > 0 && 1.0 - (r.ncc_sum + rem) / (r.weight_sum + rem) > reject_above)
      return Cost{1e30, 0.0, 0.0, 0, 0};
    const recon_mvs_camera& sc = p.cams[1 + k];
    double ps[3];
    WorldToCam(sc, pw, ps);
    if (ps[2] <= 1e-6) continue;
    double su = 0.0, sv = 0.0;
    if (!Project(sc, ps, &su, &sv)) continue;
    const Ncc n = ncc(k, su, sv);
    if (!n.valid) continue;
    ++r.valid_views;
    if (n.weak) ++r.weak;
    r.ncc_sum += 1.0 * n.ncc;
    r.weight_sum += 1.0;
  }
  if (r.weight_sum <= 0.0) return r;
  const double mean = r.ncc_sum / r.weight_sum;
  r.cost = 1.0 - mean;
  if (r.cost > 3.0 * p.photo_threshold) r.cost = 1e30;
  return r;
}

// Reference patch at any (u, v), sampled from the image.
__device__ Cost Evaluate(const RefProblem& p, double u, double v, const double* ray, double inv) {
  return EvaluateWith(p, ray, inv, [&](int k, double su, double sv) {
    return NccU8(p.ref, p.rw, p.rh, p.src[k], p.sw[k], p.sh[k], u, v, su, sv, p.patch);
  });
}

// The reference tile of a TW x TH pixel tile: the reference samples at
// (X + 0.5, Y + 0.5) for X in [x0 - half, x0 + TW + half), Y likewise, i.e.
// the samples (x + 0.5) + dx, (y + 0.5) + dy of every patch of the tile (equal
// in FP64: all terms are small integers and halves), NaN where SampleU8
// rejects them.  Dynamic LDS of RefTileBytes(patch).
struct RefTile {
  const double* t;
  int tw;        // row length: TW + 2 * half
  int lx, ly;    // this thread's pixel in the tile
  int x, y;      // and in the image
  bool active;   // inside the image
  bool full;     // every sample of the pixel's patch inside the reference
  double rs, rs2;  // with full: the patch's sum and sum of squares, in scan order
};

inline size_t RefTileBytes(int patch) {
  const int half = patch / 2;
  return sizeof(double) * (kTileW + 2 * half) * (kTileH + 2 * half);
}

// XCD-aware tile of this workgroup: consecutive tiles go to one XCD (the
// hardware deals workgroups round-robin over the 8 XCDs), so an XCD's L2
// holds the source rows of a compact region.  A bijection on [0, gridDim.x).
__device__ __forceinline__ int XcdTile() {
  const int b = blockIdx.x, n = gridDim.x;
  const int q = n >> 3, r = n & 7, j = b & 7;
  return j * q + min(j, r) + (b >> 3);
}

__device__ RefTile LoadRefTile(const RefProblem& p, double* lds) {
  const int half = p.patch / 2;
  const int tiles_x = (p.rw + kTileW - 1) / kTileW, tiles_y = (p.rh + kTileH - 1) / kTileH;
  const int t = XcdTile();
  // Blocked order: bands of kBlockH tile rows, each walked in blocks of
  // kBlockW tile columns, raster inside a block (the last band and the last
  // block of a band may be smaller).  An XCD's consecutive tiles then cover
  // a compact image region, whose windows in the sources along any epipolar
  // direction stay closer than those of whole tile rows.
  const int band = t / (tiles_x * kBlockH);
  const int u = t - band * tiles_x * kBlockH;
  const int bh = min(kBlockH, tiles_y - band * kBlockH);
  const int blk = u / (kBlockW * bh);
  const int v = u - blk * kBlockW * bh;
  const int bw = min(kBlockW, tiles_x - blk * kBlockW);
  const int x0 = (blk * kBlockW + v % bw) * kTileW, y0 = (band * kBlockH + v / bw) * kTileH;
  const int tw = kTileW + 2 * half, th = kTileH + 2 * half;
  for (int i = threadIdx.x; i < tw * th; i += kTileThreads) {
    const int X = x0 - half + i % tw;
    const int Y = y0 - half + i / tw;
    double val = 0.0;
    lds[i] = SampleU8(p.ref, p.rw, p.rh, X + 0.5, Y + 0.5, &val) ? val : __builtin_nan("");
  }
  __syncthreads();
  RefTile r;
  r.t = lds;
  r.tw = tw;
  r.lx = threadIdx.x % kTileW;
  r.ly = threadIdx.x / kTileW;
  r.x = x0 + r.lx;
  r.y = y0 + r.ly;
  r.active = r.x < p.rw && r.y < p.rh;
  // The reference sums s_r and s_r2 of NccU8 run over the reference samples
  // alone, each its own chain of additions: when every sample is valid they
  // are the same for all sources and candidates, and summed here once, in
  // the same order, they are bit-identical.
  r.full = true;
  r.rs = 0.0;
  r.rs2 = 0.0;
  for (int dy = 0; dy <= 2 * half; ++dy) {
    const double* row = lds + (r.ly + dy) * tw + r.lx;
    for (int dx = 0; dx <= 2 * half; ++dx) {
      const double vr = row[dx];
      if (vr != vr) {
        r.full = false;
        continue;
      }
      r.rs += vr;
      r.rs2 += vr * vr;
    }
  }
  return r;
}

// NccTile when every sample pair is valid: the reference patch inside the
// reference image (rt.full) and the source window inside the source.  Then
// no sample is skipped, s_r / s_r2 are the precomputed rt.rs / rt.rs2, and
// each source sample is SampleQuad's arithmetic without its bounds tests:
// a non-negative coordinate truncates to its floor.  HALF = patch / 2 when
// known at compile time, -1 otherwise.  With HALF known, a row's columns
// (x0, fx: the same on every row) are computed once and the next row's quads
// are loaded while the current row is summed.
template <int HALF>
__device__ __forceinline__ Ncc NccTileFull(const RefTile& rt, const Quad* __restrict__ q, int bw,
                                           int bh, double su, double sv, int patch) {
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<Quad*>(q), 0, bw * bh * kQuadBytes, 0x00020000);
  NccSums a;
  if constexpr (HALF >= 0) {
    constexpr int N = 2 * HALF + 1;
    int xo[N];
    double fx[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double u = su + (i - HALF);
      const int x0 = static_cast<int>(u);
      fx[i] = u - x0;
      xo[i] = x0 * kQuadBytes;
    }
    // q[i] holds column i of the current row; once summed it is reloaded
    // with column i of the next row, consumed a whole row later.
    Quad qv[N];
    {
      const int rowoff = static_cast<int>(sv + (-HALF)) * bw * kQuadBytes;
#pragma unroll
      for (int i = 0; i < N; ++i) qv[i] = LoadQuad(rsrc, rowoff + xo[i]);
    }
#pragma unroll 1
    for (int dy = -HALF; dy <= HALF; ++dy) {
      const double v = sv + dy;
      const double fy = v - static_cast<int>(v);
      // the last row reloads itself (in bounds, unused)
      const int next = static_cast<int>(sv + (dy < HALF ? dy + 1 : dy)) * bw * kQuadBytes;
      const double* trow = rt.t + (rt.ly + dy + HALF) * rt.tw + rt.lx;
#pragma unroll
      for (int i = 0; i < N; ++i) {
        const double vs = Bilinear(qv[i], fx[i], fy);
        qv[i] = LoadQuad(rsrc, next + xo[i]);
        const double vr = trow[i];
        a.s_s += vs;
        a.s_s2 += vs * vs;
        a.s_x += vr * vs;
      }
    }
  } else {
    const int half = patch / 2;
    for (int dy = -half; dy <= half; ++dy) {
      const double v = sv + dy;
      const int y0 = static_cast<int>(v);
      const double fy = v - y0;
      const int rowoff = y0 * bw * kQuadBytes;
      const double* trow = rt.t + (rt.ly + dy + half) * rt.tw + rt.lx + half;
      for (int dx = -half; dx <= half; ++dx) {
        const double u = su + dx;
        const int x0 = static_cast<int>(u);
        const Quad t = LoadQuad(rsrc, rowoff + x0 * kQuadBytes);
        const double vs = Bilinear(t, u - x0, fy);
        const double vr = trow[dx];
        a.s_s += vs;
        a.s_s2 += vs * vs;
        a.s_x += vr * vs;
      }
    }
  }
  const int n = 2 * (HALF >= 0 ? HALF : patch / 2) + 1;
  a.s_r = rt.rs;
  a.s_r2 = rt.rs2;
  a.valid = n * n;
  return FinishNcc(a, patch);
}

// NccTile for any window, HALF known: branch-free and pipelined like
// NccTileFull.  Whether a sample pair counts is separable: the reference
// sample (x + 0.5) + dx is inside iff 0 <= x + dx <= rw - 2 (likewise rows),
// the source sample su + dx by SampleQuad's test on its column (likewise
// rows), so a pair counts iff its column and its row both do.  A pair that
// does not count adds +0.0 to every (non-negative) sum, which leaves it
// unchanged, so the sums equal NccTile's bit for bit.  Coordinates are
// clamped into the source before their conversion to int; the buffer
// descriptor bounds every load (a load outside it returns 0).
template <int HALF>
__device__ __forceinline__ Ncc NccTileMasked(const RefTile& rt, int rw, int rh,
                                             const Quad* __restrict__ q, int bw, int bh,
                                             double su, double sv, int patch) {
  constexpr int N = 2 * HALF + 1;
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<Quad*>(q), 0, bw * bh * kQuadBytes, 0x00020000);
  const double umax = static_cast<double>(bw - 1), vmax = static_cast<double>(bh - 1);
  uint32_t xo[N];
  double fx[N];
  uint32_t cols = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const double u = su + (i - HALF);
    const int X = rt.x + i - HALF;
    const bool ok = !(u < 0.0 || u > umax) && X >= 0 && X <= rw - 2;
    const int x0 = static_cast<int>(fmin(fmax(u, 0.0), umax));
    fx[i] = u - x0;
    xo[i] = static_cast<uint32_t>(x0) * kQuadBytes;
    cols |= static_cast<uint32_t>(ok) << i;
  }
  const int ncols = __builtin_popcount(cols);
  const uint32_t pitch = static_cast<uint32_t>(bw) * kQuadBytes;
  auto row_off = [&](int dy) {
    return static_cast<uint32_t>(static_cast<int>(fmin(fmax(sv + dy, 0.0), vmax))) * pitch;
  };
  NccSums a;
  Quad qv[N];
  {
    const uint32_t r0 = row_off(-HALF);
#pragma unroll
    for (int i = 0; i < N; ++i) qv[i] = LoadQuad(rsrc, r0 + xo[i]);
  }
#pragma unroll 1
  for (int dy = -HALF; dy <= HALF; ++dy) {
    const double v = sv + dy;
    const int Y = rt.y + dy;
    const bool rok = !(v < 0.0 || v > vmax) && Y >= 0 && Y <= rh - 2;
    const double fy = v - static_cast<int>(fmin(fmax(v, 0.0), vmax));
    const uint32_t next = row_off(dy < HALF ? dy + 1 : dy);
    const double* trow = rt.t + (rt.ly + dy + HALF) * rt.tw + rt.lx;
    const uint32_t m = rok ? cols : 0u;
#pragma unroll
    for (int i = 0; i < N; ++i) {
      const double vs0 = Bilinear(qv[i], fx[i], fy);
      qv[i] = LoadQuad(rsrc, next + xo[i]);
      const bool ok = (m >> i) & 1u;
      const double vs = ok ? vs0 : 0.0;
      const double vr = ok ? trow[i] : 0.0;
      a.s_r += vr;
      a.s_s += vs;
      a.s_r2 += vr * vr;
      a.s_s2 += vs * vs;
      a.s_x += vr * vs;
    }
    a.valid += rok ? ncols : 0;
  }
  return FinishNcc(a, patch);
}

// Reference patch of the tile pixel, sources sampled from their quads.  The
// source window [su - half, su + half] x [sv - half, sv + half] is inside
// the source when its extreme samples are (FP64 addition is monotonic, so
// su + dx lies between su - half and su + half); a wave takes the full-window
// path when all its active lanes can, else the masked one (HALF known) or
// the per-sample tests.
template <int HALF>
__device__ __forceinline__ Cost EvaluateTile(const RefProblem& p, const RefTile& rt,
                                             const double* ray, double inv,
                                             double reject_above = 1e300) {
  return EvaluateWith(p, ray, inv, [&](int k, double su, double sv) {
    const int half = HALF >= 0 ? HALF : p.patch / 2;
    const int bw = p.sw[k], bh = p.sh[k];
    const bool full = rt.full && su + (-half) >= 0.0 && sv + (-half) >= 0.0 &&
                      su + half <= static_cast<double>(bw - 1) &&
                      sv + half <= static_cast<double>(bh - 1);
    if (__ballot(!full) == 0) return NccTileFull<HALF>(rt, p.srcq[k], bw, bh, su, sv, p.patch);
    if constexpr (HALF >= 0)
      return NccTileMasked<HALF>(rt, p.rw, p.rh, p.srcq[k], bw, bh, su, sv, p.patch);
    else
      return NccTile(rt.t, rt.tw, rt.ly, rt.lx, p.srcq[k], bw, bh, su, sv, p.patch);
  }, reject_above);
}

inline unsigned RefTiles(int w, int h) {
  return static_cast<unsigned>(((w + kTileW - 1) / kTileW) * ((h + kTileH - 1) / kTileH));
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

__global__ void k_compute_ncc(const uint8_t* __restrict__ a, int aw, int ah,
                              const uint8_t* __restrict__ b, int bw, int bh, int n,
                              const double* __restrict__ coords, int patch,
                              recon_mvs_ncc_result* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* c = coords + 4 * i;
  const Ncc r = NccU8(a, aw, ah, b, bw, bh, c[0], c[1], c[2], c[3], patch);
  recon_mvs_ncc_result o;
  o.valid = r.valid ? 1 : 0;
  o.weak_texture = r.weak ? 1 : 0;
  o.valid_samples = r.valid_samples;
  o.reserved0 = 0;
  o.ncc = r.ncc;
  out[i] = o;
}

__global__ void k_evaluate(RefProblem p, int n, const double* __restrict__ uvi,
                           recon_mvs_cost_result* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u = uvi[3 * i], v = uvi[3 * i + 1], inv = uvi[3 * i + 2];
  double ray[3];
  PixelToRay(p.cams[0], u, v, ray);
  const Cost c = Evaluate(p, u, v, ray, inv);
  recon_mvs_cost_result o;
  o.cost = c.cost;
  o.ncc_sum = c.ncc_sum;
  o.weight_sum = c.weight_sum;
  o.weak_texture_count = c.weak;
  o.valid_views = c.valid_views;
  out[i] = o;
}

// Initial cost of every pixel (patchmatch_stereo.cpp:543-581).
// One RefTile per workgroup (RefTiles(w, h) workgroups of kTileThreads);
// HALF = patch / 2, or -1 for any patch.
template <int HALF>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_pm_init_cost(RefProblem p,
                                                               const double* __restrict__ inv,
                                                               double* __restrict__ cost,
                                                               int* __restrict__ support) {
  extern __shared__ double lds[];
  const RefTile rt = LoadRefTile(p, lds);
  if (!rt.active) return;
  const int idx = rt.y * p.rw + rt.x;
  const double u = rt.x + 0.5, v = rt.y + 0.5;
  double ray[3];
  PixelToRay(p.cams[0], u, v, ray);
  const Cost c = EvaluateTile<HALF>(p, rt, ray, inv[idx]);
  cost[idx] = c.cost;
  support[idx] = c.valid_views;
}

// One directional pass (patchmatch_stereo.cpp:655-721) for every pixel.
// tab holds this pass's replayed draws, SoA [K + 2][P]: K random-search draws,
// the restart test draw, the restart value draw (0 when not drawn).
// One RefTile per workgroup (RefTiles(w, h) workgroups of kTileThreads);
// HALF = patch / 2, or -1 for any patch.
template <int HALF>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_pm_pass(RefProblem p, const double* __restrict__ snap_inv,
                          const double* __restrict__ snap_cost, const int* __restrict__ snap_sup,
                          double* __restrict__ out_inv, double* __restrict__ out_cost,
                          int* __restrict__ out_sup, const uint32_t* __restrict__ tab, int K,
                          double ratio, double restart_ratio, double inv_min, double inv_max,
                          int propagate, uint32_t* __restrict__ restart_list,
                          unsigned int* __restrict__ restart_count) {
  extern __shared__ double lds[];
  const int P = p.rw * p.rh;
  const RefTile rt = LoadRefTile(p, lds);
  if (!rt.active) return;
  const int x = rt.x, y = rt.y;
  const int idx = y * p.rw + x;
  const double u = x + 0.5, v = y + 0.5;
  double ray[3];
  PixelToRay(p.cams[0], u, v, ray);
  double best_c = snap_cost[idx];
  double best_i = snap_inv[idx];
  int best_s = snap_sup[idx];
  if (propagate) {
    const int nbx[4] = {-1, 1, 0, 0};
    const int nby[4] = {0, 0, -1, 1};
    // A candidate equal to the current best, or to a neighbour already
    // evaluated here, cannot win: its cost is the same function of (pixel,
    // inverse depth) as the one it would be compared with (the state always
    // holds the cost of its own inverse depth), and the comparison is
    // strict.  Propagation spreads identical values, so converged regions
    // skip many of these evaluations; the result is unchanged.
    double seen[4];
    int nseen = 0;
    for (int k = 0; k < 4; ++k) {
      const int nx = x + nbx[k], ny = y + nby[k];
      if (nx < 0 || ny < 0 || nx >= p.rw || ny >= p.rh) continue;
      const int nidx = ny * p.rw + nx;
      if (snap_cost[nidx] >= 1e29) continue;
      const double cand = snap_inv[nidx];
      bool dup = cand == best_i;
#pragma unroll
      for (int q = 0; q < 4; ++q) dup |= q < nseen && seen[q] == cand;
      if (dup) continue;
#pragma unroll
      for (int q = 0; q < 4; ++q)
        if (q == nseen) seen[q] = cand;
      ++nseen;
      const Cost c = EvaluateTile<HALF>(p, rt, ray, cand, best_c + kRejectMargin);
      if (c.cost < best_c) {
        best_c = c.cost;
        best_i = cand;
        best_s = c.valid_views;
      }
    }
  }
  double radius = best_i * ratio;
  for (int k = 0; k < K; ++k) {
    const double sgn = static_cast<double>(tab[static_cast<size_t>(k) * P + idx]) / 4294967296.0 * 2.0 - 1.0;
    const double cand = best_i + sgn * radius;
    if (cand < inv_min || cand > inv_max) {
      radius *= 0.5;
      continue;
    }
    const Cost c = EvaluateTile<HALF>(p, rt, ray, cand, best_c + kRejectMargin);
    if (c.cost < best_c) {
      best_c = c.cost;
      best_i = cand;
      best_s = c.valid_views;
    }
    radius *= 0.5;
  }
  out_inv[idx] = best_i;
  out_cost[idx] = best_c;
  out_sup[idx] = best_s;
  // The restart candidate (drawn by ~restart_ratio of the pixels) is the
  // pass's last step: the pixel goes to a list that k_pm_restart evaluates
  // densely, instead of whole waves running it for a few lanes.  One atomic
  // per wave reserves the wave's entries.
  const double ur = static_cast<double>(tab[static_cast<size_t>(K) * P + idx]) / 4294967296.0;
  const bool restart = ur < restart_ratio;
  const unsigned long long m = __ballot(restart);
  if (m) {
    const int lane = static_cast<int>(__lane_id());
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    unsigned int base = 0;
    if (lane == leader) base = atomicAdd(restart_count, static_cast<unsigned int>(__popcll(m)));
    base = __shfl(base, leader);
    if (restart)
      restart_list[base + static_cast<unsigned int>(__popcll(m & ((1ull << lane) - 1ull)))] =
          static_cast<uint32_t>(idx);
  }
}

// NccU8 with the source sampled from its quads (same values).
__device__ Ncc NccU8Quad(const uint8_t* __restrict__ a, int aw, int ah, const Quad* __restrict__ q,
                         int bw, int bh, double ru, double rv, double su, double sv, int patch) {
  const int half = patch / 2;
  NccSums acc;
  for (int dy = -half; dy <= half; ++dy) {
    for (int dx = -half; dx <= half; ++dx) {
      double vr = 0.0, vs = 0.0;
      if (!SampleU8(a, aw, ah, ru + dx, rv + dy, &vr)) continue;
      if (!SampleQuad(q, bw, bh, su + dx, sv + dy, &vs)) continue;
      acc.add(vr, vs);
    }
  }
  return FinishNcc(acc, patch);
}

// Lanes per listed pixel in k_pm_restart: the pixel's source views are
// evaluated in rounds of kRestartLanes, view base + j on lane j.
constexpr int kRestartLanes = 4;

// The restart step of a pass (patchmatch_stereo.cpp:707-719) for the pixels
// k_pm_pass listed, against the state it wrote: EvaluateCandidate with each
// round of views computed by a group of kRestartLanes lanes and accumulated
// in view order by every lane of the group (EvaluateWith's order).  Between
// rounds the early-rejection bound of EvaluateWith against the stored cost
// stops a candidate that cannot win (a random restart rarely can).  Any grid
// of whole groups: each group strides over the list up to the count the
// pass left.
__global__ void k_pm_restart(RefProblem p, const uint32_t* __restrict__ list,
                             const unsigned int* __restrict__ count, const uint32_t* __restrict__ tab,
                             int K, double inv_min, double inv_max, double* __restrict__ inv,
                             double* __restrict__ cost, int* __restrict__ sup) {
  static_assert(64 % kRestartLanes == 0, "groups within a wave");
  const size_t P = static_cast<size_t>(p.rw) * p.rh;
  const unsigned int n = *count;
  const unsigned int groups = gridDim.x * blockDim.x / kRestartLanes;
  const int j = static_cast<int>(threadIdx.x % kRestartLanes);
  const int leader = static_cast<int>(__lane_id()) - j;
  for (unsigned int i = (blockIdx.x * blockDim.x + threadIdx.x) / kRestartLanes; i < n; i += groups) {
    const uint32_t idx = list[i];
    const int x = static_cast<int>(idx % p.rw), y = static_cast<int>(idx / p.rw);
    const double u = x + 0.5, v = y + 0.5;
    double ray[3];
    PixelToRay(p.cams[0], u, v, ray);
    const double uv = static_cast<double>(tab[(K + 1) * P + idx]) / 4294967296.0;
    const double cand = inv_min + uv * (inv_max - inv_min);
    const double best = cost[idx];
    // EvaluateWith's opening tests and world point
    bool go = !(cand <= 1e-9);
    double pw[3] = {0.0, 0.0, 0.0};
    if (go) {
      const double depth = 1.0 / cand;
      go = !(depth < p.min_depth || depth > p.max_depth);
      if (go) {
        const double pc[3] = {ray[0] * depth, ray[1] * depth, ray[2] * depth};
        CamToWorld(p.cams[0], pc, pw);
      }
    }
    Cost r{1e30, 0.0, 0.0, 0, 0};
    bool rejected = false;
    for (int base = 0; go && base < p.nsrc; base += kRestartLanes) {
      if (base > 0) {
        const double rem = static_cast<double>(p.nsrc - base);
        if (1.0 - (r.ncc_sum + rem) / (r.weight_sum + rem) > best + kRejectMargin) {
          rejected = true;
          break;
        }
      }
      // this lane's view
      const int k = base + j;
      int valid = 0, weak = 0;
      double nv = 0.0;
      if (k < p.nsrc) {
        const recon_mvs_camera& sc = p.cams[1 + k];
        double ps[3];
        WorldToCam(sc, pw, ps);
        double su = 0.0, sv = 0.0;
        if (ps[2] > 1e-6 && Project(sc, ps, &su, &sv)) {
          const Ncc nn = NccU8Quad(p.ref, p.rw, p.rh, p.srcq[k], p.sw[k], p.sh[k], u, v, su, sv, p.patch);
          if (nn.valid) {
            valid = 1;
            weak = nn.weak ? 1 : 0;
            nv = nn.ncc;
          }
        }
      }
      // the round's views in order, on every lane of the group
      for (int jj = 0; jj < kRestartLanes; ++jj) {
        if (base + jj >= p.nsrc) break;
        const int vb = __shfl(valid, leader + jj);
        const int wb = __shfl(weak, leader + jj);
        const double nvj = __shfl(nv, leader + jj);
        if (!vb) continue;
        ++r.valid_views;
        if (wb) ++r.weak;
        r.ncc_sum += 1.0 * nvj;
        r.weight_sum += 1.0;
      }
    }
    if (j != 0 || !go || rejected) continue;
    if (r.weight_sum > 0.0) {
      const double mean = r.ncc_sum / r.weight_sum;
      r.cost = 1.0 - mean;
      if (r.cost > 3.0 * p.photo_threshold) r.cost = 1e30;
    }
    if (r.cost < best) {
      cost[idx] = r.cost;
      inv[idx] = cand;
      sup[idx] = r.valid_views;
    }
  }
}

// Output stage (patchmatch_stereo.cpp:746-763).
__global__ void k_pm_output(int P, const double* __restrict__ inv, const double* __restrict__ cost,
                            const int* __restrict__ sup, double min_depth, double max_depth,
                            float* __restrict__ depth, uint8_t* __restrict__ support,
                            float* __restrict__ conf, unsigned int* __restrict__ filled) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= P) return;
  float d_out = __builtin_nanf("");
  uint8_t s_out = 0;
  float c_out = 0.0f;
  bool ok = false;
  const double c = cost[idx];
  if (c < 1e29) {
    const double d = InvToDepth(inv[idx]);
    if (!(d < min_depth || d > max_depth)) {
      d_out = static_cast<float>(d);
      s_out = static_cast<uint8_t>(min(sup[idx], 255));
      const double pc = fmin(fmax(c, 0.0), 1.0);
      c_out = static_cast<float>(1.0 - pc);
      ok = true;
    }
  }
  depth[idx] = d_out;
  support[idx] = s_out;
  conf[idx] = c_out;
  // one atomic per wavefront
  const unsigned long long m = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(filled, static_cast<unsigned int>(__popcll(m)));
}

// Quads of an image for SampleQuad: texel (y, x) = MakeQuad(I(y, x), I(y, x1),
// I(y1, x), I(y1, x1)), x1 = min(x + 1, w - 1), y1 = min(y + 1, h - 1).
__global__ void k_make_quads(const uint8_t* __restrict__ img, int w, int h, Quad* __restrict__ quad) {
  const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<size_t>(w) * h) return;
  const int x = static_cast<int>(i % w), y = static_cast<int>(i / w);
  const int x1 = min(x + 1, w - 1), y1 = min(y + 1, h - 1);
  const uint8_t* r0 = img + static_cast<size_t>(y) * w;
  const uint8_t* r1 = img + static_cast<size_t>(y1) * w;
  quad[static_cast<size_t>(y) * w + x] = MakeQuad(r0[x], r0[x1], r1[x], r1[x1]);
}

__device__ int Solve3(double A[3][3], double b[3], double x[3]) {
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

// RefineSubpixel body for one pixel (patchmatch_stereo.cpp:807-976).
// One RefTile per workgroup (RefTiles(w, h) workgroups of kTileThreads);
// HALF = patch / 2, or -1 for any patch.
template <int HALF>
__global__ __launch_bounds__(kTileThreads) __attribute__((amdgpu_waves_per_eu(2))) void k_refine(RefProblem p, float* __restrict__ depth,
                                                         float* __restrict__ conf, double inv_min,
                                                         double inv_max, double inv_per_px,
                                                         int fit_points, int iterations,
                                                         unsigned int* __restrict__ refined) {
  extern __shared__ double lds[];
  const RefTile rt = LoadRefTile(p, lds);
  bool ok = false;
  if (rt.active) {
    const int idx = rt.y * p.rw + rt.x;
    const float d0 = depth[idx];
    const double u = rt.x + 0.5, v = rt.y + 0.5;
    const double inv0 = DepthToInv(d0);
    if (isfinite(d0) && !(inv0 < inv_min || inv0 > inv_max)) {
      double ray[3];
      PixelToRay(p.cams[0], u, v, ray);
      const int half_fit = fit_points / 2;
      const double fit_step = inv_per_px * 0.5;
      // The fit costs enter only the normal equations, which are used only
      // when every one is finite: they are summed as they come, in the
      // reference's order, and the scan stops at the first failure (the
      // reference evaluates the remaining ones and discards them).
      double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, t0 = 0, t1 = 0, t2 = 0;
      bool fit_ok = true;
      // the middle fit point is inv0 + 0.0 = inv0: its cost is the cost at
      // inv0 the reference evaluates next, reused when the scan got there
      bool have_c0 = false;
      double c0 = 0.0;
      for (int k = 0; k < fit_points; ++k) {
        const double d = inv0 + (k - half_fit) * fit_step;
        if (d < inv_min || d > inv_max) {
          fit_ok = false;
          break;
        }
        const double c = EvaluateTile<HALF>(p, rt, ray, d).cost;
        if (k == half_fit) {
          have_c0 = true;
          c0 = c;
        }
        if (c >= 1e29) {
          fit_ok = false;
          break;
        }
        const double xx = static_cast<double>(k - half_fit);
        s0 += 1.0;
        s1 += xx;
        s2 += xx * xx;
        s3 += xx * xx * xx;
        s4 += xx * xx * xx * xx;
        t0 += c;
        t1 += xx * c;
        t2 += xx * xx * c;
      }
      double best_d = inv0;
      double best_c = have_c0 ? c0 : EvaluateTile<HALF>(p, rt, ray, inv0).cost;
      if (fit_ok) {
        double A[3][3] = {{s4, s3, s2}, {s3, s2, s1}, {s2, s1, s0}};
        double b[3] = {t2, t1, t0};
        double sol[3] = {0.0, 0.0, 0.0};
        if (Solve3(A, b, sol)) {
          const double a = sol[0], bb = sol[1];
          if (fabs(a) > 1e-12) {
            const double vertex = -bb / (2.0 * a);
            if (vertex >= -static_cast<double>(half_fit) && vertex <= static_cast<double>(half_fit)) {
              // The reference adds the vertex in sample-index units (line 898).
              const double cand = inv0 + vertex;
              // compared with best_c only: rejected early when it cannot win
              const double cc = EvaluateTile<HALF>(p, rt, ray, cand, best_c + kRejectMargin).cost;
              if (cc < best_c) {
                best_c = cc;
                best_d = cand;
              }
            }
          }
        }
      }
      double lo = best_d - inv_per_px * 1.0;
      double hi = best_d + inv_per_px * 1.0;
      if (lo < inv_min) lo = inv_min;
      if (hi > inv_max) hi = inv_max;
      if (hi > lo) {
        const double inv_phi = 0.6180339887498949;
        double c1 = hi - inv_phi * (hi - lo);
        double c2 = lo + inv_phi * (hi - lo);
        double f1 = EvaluateTile<HALF>(p, rt, ray, c1).cost;
        double f2 = EvaluateTile<HALF>(p, rt, ray, c2).cost;
        for (int k = 0; k < iterations; ++k) {
          // each lane moves its bracket, then one evaluation serves both
          // sides (lanes that went different ways would otherwise run the
          // evaluation once per side); the same points in the same order
          const bool left = f1 < f2;
          double cn;
          if (left) {
            hi = c2;
            c2 = c1;
            f2 = f1;
            cn = hi - inv_phi * (hi - lo);
          } else {
            lo = c1;
            c1 = c2;
            f1 = f2;
            cn = lo + inv_phi * (hi - lo);
          }
          const double fn = EvaluateTile<HALF>(p, rt, ray, cn).cost;
          if (left) {
            c1 = cn;
            f1 = fn;
          } else {
            c2 = cn;
            f2 = fn;
          }
        }
        const double cm = 0.5 * (lo + hi);
        const double fm = EvaluateTile<HALF>(p, rt, ray, cm, best_c + kRejectMargin).cost;
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
      if (!(best_d < inv_min || best_d > inv_max)) {
        const double fd = InvToDepth(best_d);
        if (!(fd < p.min_depth || fd > p.max_depth)) {
          depth[idx] = static_cast<float>(fd);
          const double pc = fmin(fmax(best_c, 0.0), 1.0);
          conf[idx] = static_cast<float>(1.0 - pc);
          ok = true;
        }
      }
    }
  }
  const unsigned long long m = __ballot(ok);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(refined, static_cast<unsigned int>(__popcll(m)));
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

inline uint32_t XsNext(uint64_t* s) {
  *s ^= *s >> 12;
  *s ^= *s << 25;
  *s ^= *s >> 27;
  return static_cast<uint32_t>((*s * 0x2545F4914F6CDD1DULL) >> 32);
}
inline uint64_t XsSeed(uint64_t seed) { return seed ? seed : 0x9E3779B97F4A7C15ULL; }

int ResolveThreads(int n) {
  if (n > 0) return n;
  const unsigned hw = std::thread::hardware_concurrency();
  return static_cast<int>(hw > 0 ? hw : 1);
}

// Replays the row-block xorshift streams of one pass (iteration `iter`,
// direction `dir`) into tab (SoA [K + 2][P]); patchmatch_stereo.cpp:605-721.
void FillPassTable(uint32_t* tab, int w, int h, int iter, int dir, int threads, int K,
                   double restart_ratio) {
  const size_t P = static_cast<size_t>(w) * h;
  const bool rev_x = (dir == 1 || dir == 3);
  const bool rev_y = (dir >= 2);
  const int block_count = std::min(threads, std::max(1, h / 16));
  const int rows = (h + block_count - 1) / block_count;
  auto run = [&](int b) {
    const int y0 = b * rows;
    const int y1 = std::min(h, y0 + rows);
    if (y0 >= y1) return;
    uint64_t s = XsSeed(0x5DEECE66DULL + static_cast<uint64_t>(b) * 0x9E3779B97F4A7C15ULL +
                        static_cast<uint64_t>(iter) * 0xC2B2AE3D27D4EB4FULL);
    for (int yy = 0; yy < y1 - y0; ++yy) {
      const int y = rev_y ? y1 - 1 - yy : y0 + yy;
      for (int xx = 0; xx < w; ++xx) {
        const int x = rev_x ? w - 1 - xx : xx;
        const size_t idx = static_cast<size_t>(y) * w + x;
        for (int k = 0; k < K; ++k) tab[k * P + idx] = XsNext(&s);
        const uint32_t r = XsNext(&s);
        tab[K * P + idx] = r;
        tab[(K + 1) * P + idx] =
            (static_cast<double>(r) / 4294967296.0 < restart_ratio) ? XsNext(&s) : 0u;
      }
    }
  };
  const int host_threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  if (block_count <= 1 || host_threads <= 1) {
    for (int b = 0; b < block_count; ++b) run(b);
  } else {
    std::vector<std::thread> ws;
    const int nw = std::min(block_count, host_threads);
    for (int t = 0; t < nw; ++t)
      ws.emplace_back([&, t]() {
        for (int b = t; b < block_count; b += nw) run(b);
      });
    for (auto& th : ws) th.join();
  }
}

int32_t ValidateProblem(recon_mvs_hip_context* ctx, const recon_mvs_problem* pr,
                        const recon_mvs_config* cfg) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!pr || !cfg) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null problem or config");
  if (pr->pixel_format != RECON_MVS_PIXEL_U8)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "reference mode takes uint8 images");
  if (!pr->ref_pixels || pr->ref_w <= 0 || pr->ref_h <= 0)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad reference image");
  if (pr->num_src < 0 || pr->num_src > kMaxSrc)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "num_src out of [0, 32]");
  if (pr->num_src > 0 && (!pr->src_pixels || !pr->src_w || !pr->src_h || !pr->src_cameras))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null source arrays");
  for (int k = 0; k < pr->num_src; ++k)
    if (!pr->src_pixels[k] || pr->src_w[k] <= 0 || pr->src_h[k] <= 0)
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad source image");
  // pixel indices and the quads' byte offsets are 32-bit
  constexpr int64_t kMaxPixels = int64_t{1} << 29;
  if (static_cast<int64_t>(pr->ref_w) * pr->ref_h > kMaxPixels)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "reference image above 2^29 pixels");
  for (int k = 0; k < pr->num_src; ++k)
    if (static_cast<int64_t>(pr->src_w[k]) * pr->src_h[k] > kMaxPixels)
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "source image above 2^29 pixels");
  if (cfg->patch_size < 1 || cfg->patch_size > 63)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "patch_size out of [1, 63]");
  return RECON_MVS_OK;
}

inline unsigned Blocks(size_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

// Calls f(std::integral_constant<int, HALF>) with the pixel kernels'
// instantiation for the patch: unrolled for 7x7 and 11x11, generic otherwise.
template <class F>
void WithHalf(int patch, F f) {
  switch (patch / 2) {
    case 3: f(std::integral_constant<int, 3>()); break;
    case 5: f(std::integral_constant<int, 5>()); break;
    default: f(std::integral_constant<int, -1>()); break;
  }
}

// Uploads images and cameras of a problem; fills the device RefProblem.  With
// `quads`, also builds each source's quad image (for the tile kernels).
int32_t Upload(recon_mvs_hip_context* ctx, const recon_mvs_problem* pr,
               const recon_mvs_config* cfg, RefProblem* dp, bool quads = false) {
  size_t total = static_cast<size_t>(pr->ref_w) * pr->ref_h;
  std::vector<size_t> off(pr->num_src);
  for (int k = 0; k < pr->num_src; ++k) {
    total = (total + 255) & ~static_cast<size_t>(255);
    off[k] = total;
    total += static_cast<size_t>(pr->src_w[k]) * pr->src_h[k];
  }
  uint8_t* img = static_cast<uint8_t*>(Ensure(ctx, kSlotRefImg, total));
  auto* cams = static_cast<recon_mvs_camera*>(
      Ensure(ctx, kSlotCams, sizeof(recon_mvs_camera) * (1 + pr->num_src)));
  if (!img || !cams) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "device allocation failed");
  RMVS_HIP(ctx, hipMemcpyAsync(img, pr->ref_pixels, static_cast<size_t>(pr->ref_w) * pr->ref_h,
                               hipMemcpyHostToDevice, ctx->stream));
  for (int k = 0; k < pr->num_src; ++k)
    RMVS_HIP(ctx, hipMemcpyAsync(img + off[k], pr->src_pixels[k],
                                 static_cast<size_t>(pr->src_w[k]) * pr->src_h[k],
                                 hipMemcpyHostToDevice, ctx->stream));
  std::vector<recon_mvs_camera> hc(1 + pr->num_src);
  hc[0] = pr->ref_camera;
  for (int k = 0; k < pr->num_src; ++k) hc[1 + k] = pr->src_cameras[k];
  RMVS_HIP(ctx, hipMemcpyAsync(cams, hc.data(), sizeof(recon_mvs_camera) * hc.size(),
                               hipMemcpyHostToDevice, ctx->stream));
  // the host vector dies here: make the copy complete first
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  std::memset(dp, 0, sizeof(*dp));
  dp->ref = img;
  dp->rw = pr->ref_w;
  dp->rh = pr->ref_h;
  dp->nsrc = pr->num_src;
  dp->patch = cfg->patch_size;
  for (int k = 0; k < pr->num_src; ++k) {
    dp->src[k] = img + off[k];
    dp->sw[k] = pr->src_w[k];
    dp->sh[k] = pr->src_h[k];
  }
  if (quads) {
    std::vector<size_t> qoff(pr->num_src);
    size_t qtotal = 0;
    for (int k = 0; k < pr->num_src; ++k) {
      qoff[k] = qtotal;
      qtotal += ((static_cast<size_t>(pr->src_w[k]) * pr->src_h[k] * kQuadBytes + 255) & ~static_cast<size_t>(255));
    }
    auto* q = static_cast<uint8_t*>(Ensure(ctx, kSlotSrcImg, qtotal));
    if (!q) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "device allocation failed");
    for (int k = 0; k < pr->num_src; ++k) {
      auto* qk = reinterpret_cast<Quad*>(q + qoff[k]);
      hipLaunchKernelGGL(k_make_quads, dim3(Blocks(static_cast<size_t>(pr->src_w[k]) * pr->src_h[k], 256)),
                         dim3(256), 0,
                         ctx->stream, img + off[k], pr->src_w[k], pr->src_h[k], qk);
      RMVS_LAUNCH_CHECK(ctx);
      dp->srcq[k] = qk;
    }
  }
  dp->cams = cams;
  dp->min_depth = cfg->min_depth_m;
  dp->max_depth = cfg->max_depth_m;
  dp->photo_threshold = cfg->photo_threshold;
  return RECON_MVS_OK;
}

}  // namespace
}  // namespace rmvs

using namespace rmvs;

extern "C" {

RECON_MVS_API int32_t recon_mvs_hip_compute_ncc(recon_mvs_hip_context* ctx, const uint8_t* ref_img,
                                                int32_t ref_w, int32_t ref_h,
                                                const uint8_t* src_img, int32_t src_w,
                                                int32_t src_h, int32_t n, const double* coords,
                                                int32_t patch_size, recon_mvs_ncc_result* out) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!ref_img || !src_img || ref_w <= 0 || ref_h <= 0 || src_w <= 0 || src_h <= 0 || n < 0 ||
      (n > 0 && (!coords || !out)) || patch_size < 1 || patch_size > 63)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "compute_ncc: bad arguments");
  if (n == 0) return RECON_MVS_OK;
  const size_t ra = static_cast<size_t>(ref_w) * ref_h, sa = static_cast<size_t>(src_w) * src_h;
  auto* da = static_cast<uint8_t*>(Ensure(ctx, kSlotQueryImgA, ra));
  auto* db = static_cast<uint8_t*>(Ensure(ctx, kSlotQueryImgB, sa));
  auto* dc = static_cast<double*>(Ensure(ctx, kSlotQueryIn, sizeof(double) * 4 * n));
  auto* dout = static_cast<recon_mvs_ncc_result*>(
      Ensure(ctx, kSlotQueryOut, sizeof(recon_mvs_ncc_result) * n));
  if (!da || !db || !dc || !dout) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  RMVS_HIP(ctx, hipMemcpyAsync(da, ref_img, ra, hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(db, src_img, sa, hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(dc, coords, sizeof(double) * 4 * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_compute_ncc, dim3(Blocks(n, 256)), dim3(256), 0, ctx->stream, da, ref_w,
                     ref_h, db, src_w, src_h, n, dc, patch_size, dout);
  RMVS_LAUNCH_CHECK(ctx);
  RMVS_HIP(ctx, hipMemcpyAsync(out, dout, sizeof(recon_mvs_ncc_result) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_evaluate_candidates(recon_mvs_hip_context* ctx,
                                                        const recon_mvs_problem* problem,
                                                        const recon_mvs_config* config, int32_t n,
                                                        const double* u_v_inv,
                                                        recon_mvs_cost_result* out) {
  int32_t st = ValidateProblem(ctx, problem, config);
  if (st != RECON_MVS_OK) return st;
  if (n < 0 || (n > 0 && (!u_v_inv || !out)))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "evaluate_candidates: bad query arrays");
  if (n == 0) return RECON_MVS_OK;
  RefProblem dp;
  st = Upload(ctx, problem, config, &dp);
  if (st != RECON_MVS_OK) return st;
  auto* dq = static_cast<double*>(Ensure(ctx, kSlotQueryIn, sizeof(double) * 3 * n));
  auto* dout = static_cast<recon_mvs_cost_result*>(
      Ensure(ctx, kSlotQueryOut, sizeof(recon_mvs_cost_result) * n));
  if (!dq || !dout) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  RMVS_HIP(ctx, hipMemcpyAsync(dq, u_v_inv, sizeof(double) * 3 * n, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_evaluate, dim3(Blocks(n, 128)), dim3(128), 0, ctx->stream, dp, n, dq, dout);
  RMVS_LAUNCH_CHECK(ctx);
  RMVS_HIP(ctx, hipMemcpyAsync(out, dout, sizeof(recon_mvs_cost_result) * n, hipMemcpyDeviceToHost,
                               ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_patchmatch_stereo(recon_mvs_hip_context* ctx,
                                                      const recon_mvs_problem* problem,
                                                      const recon_mvs_config* config,
                                                      const double* seed_points, int32_t num_seeds,
                                                      float* depth, uint8_t* support,
                                                      float* confidence, uint32_t* filled) {
  int32_t st = ValidateProblem(ctx, problem, config);
  if (st != RECON_MVS_OK) return st;
  if (!depth || !support || !confidence || !filled)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null output");
  if (num_seeds < 0 || (num_seeds > 0 && !seed_points))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad seed points");
  if (config->random_search_count < 0 || config->random_search_count > 64)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "random_search_count out of [0, 64]");
  *filled = 0;
  // source_indices.empty() -> 0 before DepthMap::Allocate (lines 434-435)
  if (problem->num_src == 0) return RECON_MVS_OK;
  const int w = problem->ref_w, h = problem->ref_h;
  const size_t P = static_cast<size_t>(w) * h;
  for (size_t i = 0; i < P; ++i) {  // DepthMap::Allocate (mvs.h:95-102)
    depth[i] = __builtin_nanf("");
    support[i] = 0;
    confidence[i] = 0.0f;
  }
  if (MeanBaseline(problem->ref_camera, problem->src_cameras, problem->num_src) < 1e-9)
    return RECON_MVS_OK;
  const double inv_min = DepthToInv(config->max_depth_m);
  const double inv_max = DepthToInv(config->min_depth_m);
  if (!(inv_max > inv_min) || inv_min <= 0.0) return RECON_MVS_OK;

  RefProblem dp;
  st = Upload(ctx, problem, config, &dp, true);
  if (st != RECON_MVS_OK) return st;

  // Seeding + random init on the host: one sequential stream (lines 496-539).
  std::vector<double> inv0(P, 0.0);
  std::vector<uint8_t> init(P, 0);
  if (seed_points && num_seeds > 0) {
    const int grid = 32;
    const int gx = (w + grid - 1) / grid, gy = (h + grid - 1) / grid;
    std::vector<std::vector<double>> cell(static_cast<size_t>(gx) * gy);
    for (int i = 0; i < num_seeds; ++i) {
      const double su = seed_points[3 * i], sv = seed_points[3 * i + 1], sd = seed_points[3 * i + 2];
      if (su < 0 || sv < 0 || su >= w || sv >= h) continue;
      if (sd <= 0.0) continue;
      const int cx = std::min(gx - 1, static_cast<int>(su) / grid);
      const int cy = std::min(gy - 1, static_cast<int>(sv) / grid);
      cell[static_cast<size_t>(cy) * gx + cx].push_back(sd);
    }
    std::vector<double> mean(cell.size(), 0.0);
    for (size_t c = 0; c < cell.size(); ++c) {
      if (cell[c].empty()) continue;
      double s = 0.0;
      for (double d : cell[c]) s += d;
      mean[c] = s / static_cast<double>(cell[c].size());
    }
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x) {
        const size_t c = static_cast<size_t>(std::min(gy - 1, y / grid)) * gx + std::min(gx - 1, x / grid);
        if (cell[c].empty()) continue;
        inv0[static_cast<size_t>(y) * w + x] = DepthToInv(mean[c]);
        init[static_cast<size_t>(y) * w + x] = 1;
      }
  }
  {
    uint64_t s = XsSeed(0x5DEECE66DULL);
    for (size_t i = 0; i < P; ++i) {
      if (init[i]) continue;
      inv0[i] = inv_min + (static_cast<double>(XsNext(&s)) / 4294967296.0) * (inv_max - inv_min);
    }
  }

  double* invA = static_cast<double*>(Ensure(ctx, kSlotInvA, P * sizeof(double)));
  double* invB = static_cast<double*>(Ensure(ctx, kSlotInvB, P * sizeof(double)));
  double* costA = static_cast<double*>(Ensure(ctx, kSlotCostA, P * sizeof(double)));
  double* costB = static_cast<double*>(Ensure(ctx, kSlotCostB, P * sizeof(double)));
  int* supA = static_cast<int*>(Ensure(ctx, kSlotSupA, P * sizeof(int)));
  int* supB = static_cast<int*>(Ensure(ctx, kSlotSupB, P * sizeof(int)));
  const int K = config->random_search_count;
  const size_t tab_bytes = static_cast<size_t>(K + 2) * P * sizeof(uint32_t);
  uint32_t* tab[2] = {static_cast<uint32_t*>(Ensure(ctx, kSlotRng0, tab_bytes)),
                      static_cast<uint32_t*>(Ensure(ctx, kSlotRng1, tab_bytes))};
  float* d_depth = static_cast<float*>(Ensure(ctx, kSlotOutDepth, P * sizeof(float)));
  uint8_t* d_sup = static_cast<uint8_t*>(Ensure(ctx, kSlotOutSupport, P));
  float* d_conf = static_cast<float*>(Ensure(ctx, kSlotOutConf, P * sizeof(float)));
  unsigned int* d_cnt = static_cast<unsigned int*>(Ensure(ctx, kSlotCounter, 64));
  // restart list of a pass: its count, then up to P pixel indices
  auto* restart = static_cast<uint32_t*>(Ensure(ctx, kSlotRestart, (P + 64) * sizeof(uint32_t)));
  if (!invA || !invB || !costA || !costB || !supA || !supB || !tab[0] || !tab[1] || !d_depth ||
      !d_sup || !d_conf || !d_cnt || !restart)
    return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "device allocation failed");
  unsigned int* restart_count = restart;
  uint32_t* restart_list = restart + 64;
  int restart_blocks = 0;
  RMVS_HIP(ctx, hipDeviceGetAttribute(&restart_blocks, hipDeviceAttributeMultiprocessorCount, ctx->device));
  restart_blocks = std::max(1, restart_blocks) * 4;

  RMVS_HIP(ctx, hipMemcpyAsync(invA, inv0.data(), P * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  const unsigned tiles = RefTiles(w, h);
  const size_t tile_lds = RefTileBytes(config->patch_size);
  hipEvent_t eb, ee;
  TimerBegin(ctx, kTimerInit, &eb, &ee);
  WithHalf(config->patch_size, [&](auto hc) {
    hipLaunchKernelGGL(k_pm_init_cost<decltype(hc)::value>, dim3(tiles), dim3(kTileThreads), tile_lds,
                       ctx->stream, dp, invA, costA, supA);
  });
  RMVS_LAUNCH_CHECK(ctx);
  TimerEnd(ctx, kTimerInit, eb, ee);

  const int threads = ResolveThreads(config->num_threads);
  const int dirs = std::max(1, config->propagation_directions);
  int pass = 0;
  for (int iter = 0; iter < config->iterations; ++iter) {
    for (int dir = 0; dir < dirs; ++dir, ++pass) {
      if (CancelRequested(ctx)) return Cancelled(ctx);
      const int b = pass & 1;
      uint32_t* host = static_cast<uint32_t*>(EnsurePinned(ctx, b, tab_bytes));
      if (!host) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "pinned allocation failed");
      // the previous upload from this staging buffer must have finished
      RMVS_HIP(ctx, hipEventSynchronize(ctx->pinned_ready[b]));
      FillPassTable(host, w, h, iter, dir, threads, K, config->restart_ratio);
      // the upload runs on the copy stream, overlapping the previous pass;
      // it waits for the pass that last read device table b, and this pass
      // waits for it
      RMVS_HIP(ctx, hipStreamWaitEvent(ctx->copy_stream, ctx->tab_free[b], 0));
      RMVS_HIP(ctx, hipMemcpyAsync(tab[b], host, tab_bytes, hipMemcpyHostToDevice, ctx->copy_stream));
      RMVS_HIP(ctx, hipEventRecord(ctx->pinned_ready[b], ctx->copy_stream));
      RMVS_HIP(ctx, hipStreamWaitEvent(ctx->stream, ctx->pinned_ready[b], 0));
      TimerBegin(ctx, kTimerPass, &eb, &ee);
      RMVS_HIP(ctx, hipMemsetAsync(restart_count, 0, sizeof(unsigned int), ctx->stream));
      WithHalf(config->patch_size, [&](auto hc) {
        hipLaunchKernelGGL(k_pm_pass<decltype(hc)::value>, dim3(tiles), dim3(kTileThreads), tile_lds,
                           ctx->stream, dp, invA, costA, supA, invB, costB, supB, tab[b], K,
                           config->random_search_radius_ratio, config->restart_ratio, inv_min,
                           inv_max, config->propagation_enabled ? 1 : 0, restart_list,
                           restart_count);
      });
      RMVS_LAUNCH_CHECK(ctx);
      hipLaunchKernelGGL(k_pm_restart, dim3(restart_blocks), dim3(256), 0, ctx->stream, dp, restart_list,
                         restart_count, tab[b], K, inv_min, inv_max, invB, costB, supB);
      RMVS_LAUNCH_CHECK(ctx);
      TimerEnd(ctx, kTimerPass, eb, ee);
      RMVS_HIP(ctx, hipEventRecord(ctx->tab_free[b], ctx->stream));
      std::swap(invA, invB);
      std::swap(costA, costB);
      std::swap(supA, supB);
    }
  }
  RMVS_HIP(ctx, hipMemsetAsync(d_cnt, 0, sizeof(unsigned int), ctx->stream));
  hipLaunchKernelGGL(k_pm_output, dim3(Blocks(P, 256)), dim3(256), 0, ctx->stream, static_cast<int>(P),
                     invA, costA, supA, config->min_depth_m, config->max_depth_m, d_depth, d_sup,
                     d_conf, d_cnt);
  RMVS_LAUNCH_CHECK(ctx);
  unsigned int cnt = 0;
  RMVS_HIP(ctx, hipMemcpyAsync(depth, d_depth, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(support, d_sup, P, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(confidence, d_conf, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(&cnt, d_cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *filled = cnt;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_refine_subpixel(recon_mvs_hip_context* ctx,
                                                    const recon_mvs_problem* problem,
                                                    const recon_mvs_config* config, float* depth,
                                                    float* confidence, uint32_t* refined) {
  int32_t st = ValidateProblem(ctx, problem, config);
  if (st != RECON_MVS_OK) return st;
  if (CancelRequested(ctx)) return Cancelled(ctx);
  if (!depth || !confidence || !refined) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null output");
  const int fit_points = std::max(3, config->subpixel_fit_points);
  if (fit_points > kMaxFitPoints)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "subpixel_fit_points above 31");
  *refined = 0;
  if (problem->num_src == 0) return RECON_MVS_OK;  // lines 780-781
  const double baseline = MeanBaseline(problem->ref_camera, problem->src_cameras, problem->num_src);
  if (baseline < 1e-9) return RECON_MVS_OK;
  const double inv_min = DepthToInv(config->max_depth_m);
  const double inv_max = DepthToInv(config->min_depth_m);
  const double focal = 0.5 * (problem->ref_camera.fx + problem->ref_camera.fy);
  const double inv_per_px = 1.0 / (focal * baseline);
  RefProblem dp;
  st = Upload(ctx, problem, config, &dp, true);
  if (st != RECON_MVS_OK) return st;
  const size_t P = static_cast<size_t>(problem->ref_w) * problem->ref_h;
  float* d_depth = static_cast<float*>(Ensure(ctx, kSlotOutDepth, P * sizeof(float)));
  float* d_conf = static_cast<float*>(Ensure(ctx, kSlotOutConf, P * sizeof(float)));
  unsigned int* d_cnt = static_cast<unsigned int*>(Ensure(ctx, kSlotCounter, 64));
  if (!d_depth || !d_conf || !d_cnt) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  RMVS_HIP(ctx, hipMemcpyAsync(d_depth, depth, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(d_conf, confidence, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  RMVS_HIP(ctx, hipMemsetAsync(d_cnt, 0, sizeof(unsigned int), ctx->stream));
  hipEvent_t eb, ee;
  TimerBegin(ctx, kTimerRefine, &eb, &ee);
  WithHalf(config->patch_size, [&](auto hc) {
    hipLaunchKernelGGL(k_refine<decltype(hc)::value>, dim3(RefTiles(problem->ref_w, problem->ref_h)),
                       dim3(kTileThreads), RefTileBytes(config->patch_size), ctx->stream, dp, d_depth,
                       d_conf, inv_min, inv_max, inv_per_px, fit_points, config->subpixel_iterations,
                       d_cnt);
  });
  RMVS_LAUNCH_CHECK(ctx);
  TimerEnd(ctx, kTimerRefine, eb, ee);
  unsigned int cnt = 0;
  RMVS_HIP(ctx, hipMemcpyAsync(depth, d_depth, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(confidence, d_conf, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(&cnt, d_cnt, sizeof(unsigned int), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *refined = cnt;
  return RECON_MVS_OK;
}

}  // extern "C"
