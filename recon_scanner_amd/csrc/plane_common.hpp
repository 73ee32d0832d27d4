// plane_common.hpp — device pieces of plane mode shared by the sweep kernels
// (plane_mode.hip) and the adjoint kernels (plane_adjoint.hip): the resident
// view and depth-map slots, the counter RNG, the kernel arguments, pixel
// access, the LDS reference tile and window, and the FP32 plane cost.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "mvs_common.hpp"

constexpr int kMaxViews = 16;
constexpr int kMaxTopK = 16;
// Tile schedules (pixels per workgroup tile, 256 threads): 0 = 64 x 8 (the
// default), 1 = 32 x 16, 2 = 128 x 4.  A wave covers 64 / (TW / 2) rows.
constexpr int kNumSchedules = 3;
constexpr float kCostInf = 1e30f;

struct ResidentView {
  int w = 0, h = 0, pitch = 0;
  int format = RECON_MVS_PIXEL_F32;
  void* pixels = nullptr;   // h rows x pitch row-pair elements (see BufQuad)
  recon_mvs_camera cam{};
};

struct PlaneState {
  int w = 0, h = 0;
  float4* plane = nullptr;  // (depth, nx, ny, nz)
  float4* prev = nullptr;   // the propagated candidates of each pixel's last half-sweep (4), by colour
  float4* cand = nullptr;   // binned half-sweeps' perturbed and random candidates: 4 per pixel of a colour
  float* cost = nullptr;
  uint8_t* support = nullptr;
  uint8_t* geo_support = nullptr;
  bool has_geo = false;
  // sparse-point seeding: per 32 x 32 pixel cell the mean seed depth (0 = no
  // seed), device copy and its host staging (kept until the copy has run)
  float* seed_cells = nullptr;
  int seed_capacity = 0;
  std::vector<float> seed_host;
};

namespace rmvs {

// ---------------------------------------------------------------------------
// Counter RNG: lowbias32 (shaders/common/gpu_algo.h:150-154, gpu_common.h:239-244)
// ---------------------------------------------------------------------------
RMVS_HD uint32_t Lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}
RMVS_HD uint32_t RngKey(uint32_t seed, uint32_t pass, uint32_t pixel) {
  return Lowbias32(Lowbias32(Lowbias32(seed) ^ pass) ^ pixel);
}
// draw k of a key, uniform in [0, 1) with 24 random bits (exact in float)
RMVS_HD float RngUniform(uint32_t key, uint32_t k) {
  return static_cast<float>(Lowbias32(key ^ ((k + 1u) * 0x9E3779B9u)) >> 8) * (1.0f / 16777216.0f);
}

// ---------------------------------------------------------------------------
// Kernel arguments
// ---------------------------------------------------------------------------
struct PlaneArgs {
  int w, h, ref_pitch;
  const void* ref_img;
  double fx, fy, cx, cy;        // reference intrinsics
  double kinv[4];               // 1/fx, 1/fy, -cx/fx, -cy/fy: the terms of K_r^-1
  float ffx, ffy, fcx, fcy;     // same, FP32
  int nsrc, top_k;
  const void* src_img[kMaxViews];
  int src_pitch[kMaxViews], src_w[kMaxViews], src_h[kMaxViews];
  // the last sampler coordinate of each source: w - 1, h - 1, and both
  // less the 1e-4 px margin of ViewNcc's inside test
  double src_lim[kMaxViews][4];
  double A[kMaxViews][9];       // K_s R_rel K_r^-1 (row-major)
  double b[kMaxViews][3];       // K_s t_rel
  float4* plane;
  // per pixel 4 planes it has been scored on in an earlier half-sweep of the
  // run (its propagated candidates; NaN: none), or null: such a plane cannot
  // beat the pixel's cost, which only falls during a run
  float4* prev;
  // 4 slots per pixel of the active colour for the perturbed and random
  // candidates (4-7) of a binned half-sweep: written by the pixel's owning
  // thread, read by the thread that scores them (same workgroup); null
  // where no binned half-sweep runs
  float4* cand;
  float* cost;
  uint8_t* support;
  const float* seed_cells;      // init: mean seed depth per 32 x 32 cell (0 = none), or null
  int seed_gx, seed_gy;
  float min_depth, max_depth, inv_min, inv_max;
  float eps2;                   // (NCC denominator floor)^2
  float photo_reject;           // 3 x photo_threshold
  uint32_t seed;
  int color, iter, reach;
  float depth_perturb, normal_perturb;
  int tiles_x, tiles_y, tiles_per_xcd, num_tiles;
};

// Resident images are stored as row pairs: element (y, x) holds
// (I[y][x], I[y+1][x]) (zero beyond the image), so the four bilinear taps of
// a sample at (x, y) are the 2 x 2 elements (y, x), (y, x + 1): one load of
// 16 B (FP32), 8 B (FP16 / BF16) or 4 B (8-bit, RECON_MVS_PIXEL_U8 storage:
// exact for 8-bit views; its 4-byte loads are 2-byte aligned).  A stride-2 window touches a dense 12 x 12
// pixel block; as row pairs that is 6 runs of 12 elements instead of 12 runs
// of 12 pixels, about half the cache lines for scattered (unconverged)
// hypotheses.  Twice the bytes of a plain image.
template <int FMT>
constexpr int ElemBytes() {
  return FMT == RECON_MVS_PIXEL_F32 ? 8 : FMT == RECON_MVS_PIXEL_U8 ? 2 : 4;
}

// Pixel (y, x) of a row-pair image, i = y * pitch + x.
template <int FMT>
__device__ __forceinline__ float Pix(const void* base, int i) {
  if constexpr (FMT == RECON_MVS_PIXEL_F32) {
    return static_cast<const float*>(base)[2 * i];
  } else if constexpr (FMT == RECON_MVS_PIXEL_U8) {
    return static_cast<float>(static_cast<const uint8_t*>(base)[2 * i]);
  } else if constexpr (FMT == RECON_MVS_PIXEL_F16) {
    return static_cast<float>(static_cast<const _Float16*>(base)[2 * i]);
  } else {
    return __uint_as_float(static_cast<uint32_t>(static_cast<const uint16_t*>(base)[2 * i]) << 16);
  }
}

// The taps of the element at byte offset `off` and its right neighbour,
// through a raw buffer descriptor (32-bit offsets; reads beyond the
// descriptor's size return zeros): (top-left, bottom-left, top-right,
// bottom-right).
template <int FMT>
__device__ __forceinline__ float4 BufQuad(__amdgpu_buffer_rsrc_t rsrc, int off) {
  if constexpr (FMT == RECON_MVS_PIXEL_F32) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rsrc, off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
  } else if constexpr (FMT == RECON_MVS_PIXEL_U8) {
    const uint32_t u = __builtin_amdgcn_raw_buffer_load_b32(rsrc, off, 0, 0);
    return make_float4(static_cast<float>(u & 0xFFu), static_cast<float>((u >> 8) & 0xFFu),
                       static_cast<float>((u >> 16) & 0xFFu), static_cast<float>(u >> 24));
  } else {
    const auto u = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0);
    if constexpr (FMT == RECON_MVS_PIXEL_F16) {
      const auto h = [](uint32_t b) {
        return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(b)));
      };
      return make_float4(h(u[0] & 0xFFFFu), h(u[0] >> 16), h(u[1] & 0xFFFFu), h(u[1] >> 16));
    } else {
      return make_float4(__uint_as_float(u[0] << 16), __uint_as_float(u[0] & 0xFFFF0000u),
                         __uint_as_float(u[1] << 16), __uint_as_float(u[1] & 0xFFFF0000u));
    }
  }
}

// Reference window of one pixel: a pointer to its top-left sample in LDS
// (row step rs floats, column step STRIDE), NaN marking samples outside the
// image; the count and mean of the samples inside, and their sums centred on
// that mean (centred sums keep the FP32 variance free of cancellation).
template <int HALF, int STRIDE>
struct RefWindow {
  static constexpr int NS = 2 * HALF / STRIDE + 1;
  static constexpr int S = NS * NS;
  const float* base;  // LDS values are pre-shifted by `shift`
  int rs;
  float sum_c, sum_c2, n, mean, shift;
  bool full;
  __device__ __forceinline__ float at(int i, int j) const { return base[j * rs + i * STRIDE]; }
};

template <int HALF, int STRIDE>
__device__ __forceinline__ void MakeRefWindow(RefWindow<HALF, STRIDE>& win, const float* base, int rs,
                                              float shift) {
  constexpr int NS = RefWindow<HALF, STRIDE>::NS;
  win.base = base;
  win.rs = rs;
  win.shift = shift;
  float s = 0.0f, cnt = 0.0f;
  for (int j = 0; j < NS; ++j) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const float v = win.at(i, j);
      const bool ok = (v == v);
      s += ok ? v : 0.0f;
      cnt += ok ? 1.0f : 0.0f;
    }
  }
  const float mean = cnt > 0.0f ? s / cnt : 0.0f;
  float sc = 0.0f, sc2 = 0.0f;
  for (int j = 0; j < NS; ++j) {
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const float v0 = win.at(i, j);
      const float c = (v0 == v0) ? v0 - mean : 0.0f;
      sc += c;
      sc2 = fmaf(c, c, sc2);
    }
  }
  win.sum_c = sc;
  win.sum_c2 = sc2;
  win.n = cnt;
  win.mean = mean;
  win.full = (cnt == static_cast<float>(RefWindow<HALF, STRIDE>::S));
}

// ---------------------------------------------------------------------------
// ALU helpers for the window loops: single instructions the compiler does not
// select from plain code (written as one-instruction asm with register
// operands only; no memory access)
// ---------------------------------------------------------------------------

// floor(x) as an integer (v_cvt_flr_i32_f32; the compiler would emit
// v_floor_f32 + v_cvt_i32_f32).
__device__ __forceinline__ int FloorInt(float x) {
  int r;
  asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// a * b + c for a, b in signed 24 bits (v_mad_i32_i24, full rate; a 32-bit
// multiply is the quarter-rate v_mul_lo_u32).
__device__ __forceinline__ int MadI24(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// The two horizontal interpolations of an FP16 row-pair quad, read straight
// from the packed halves by mixed-precision FMAs (v_fma_mix_f32: FP16 inputs
// convert exactly, each FMA rounds once in FP32).  u0 = (top-left, bottom-left),
// u1 = (top-right, bottom-right); top = x00 + fx (x10 - x00), bot likewise:
// the values FP32 arithmetic on the converted taps gives, in 4 instructions
// instead of 8.
__device__ __forceinline__ void F16Lerps(uint32_t u0, uint32_t u1, float fx, float* top, float* bot) {
  float d0, d1, t, b;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,1]" : "=v"(d0) : "v"(u0), "v"(u1));
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,1] op_sel_hi:[1,0,1]" : "=v"(d1) : "v"(u0), "v"(u1));
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel_hi:[0,0,1]" : "=v"(t) : "v"(fx), "v"(d0), "v"(u0));
  asm("v_fma_mix_f32 %0, %1, %2, %3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "=v"(b) : "v"(fx), "v"(d1), "v"(u0));
  *top = t;
  *bot = b;
}

// Raw texels of one sample (see BufQuad): FP16 stays packed for F16Lerps.
template <int FMT>
struct Texel {
  float4 q;
};
template <>
struct Texel<RECON_MVS_PIXEL_F16> {
  uint32_t u0, u1;
};

template <int FMT>
__device__ __forceinline__ Texel<FMT> LoadTexel(__amdgpu_buffer_rsrc_t rsrc, int off) {
  Texel<FMT> t;
  if constexpr (FMT == RECON_MVS_PIXEL_F16) {
    const auto u = __builtin_amdgcn_raw_buffer_load_b64(rsrc, off, 0, 0);
    t.u0 = u[0];
    t.u1 = u[1];
  } else {
    t.q = BufQuad<FMT>(rsrc, off);
  }
  return t;
}

// Bilinear value at fractions (fx, fy) of a texel quad: the same operations
// (and roundings) in every storage format.
template <int FMT>
__device__ __forceinline__ float Bilinear(const Texel<FMT>& t, float fx, float fy) {
  float top, bot;
  if constexpr (FMT == RECON_MVS_PIXEL_F16) {
    F16Lerps(t.u0, t.u1, fx, &top, &bot);
  } else {
    top = fmaf(fx, t.q.z - t.q.x, t.q.x);
    bot = fmaf(fx, t.q.w - t.q.y, t.q.y);
  }
  return fmaf(fy, bot - top, top);
}

struct Hyp {
  float d, nx, ny, nz;
};

constexpr int kMaxTopKNet = 8;  // top_k below the view count is limited to 8

// Aggregated plane cost of hypothesis h at pixel (x, y).  Returns kCostInf
// when no view is valid or the hypothesis is rejected.  view_ncc (optional)
// receives NaN for invalid views.
//
// Early exit: after each view the cost is bounded below by its value when
// every view still to come is valid with NCC = 1 (cost 0) -- adding a valid
// zero-cost view never raises the mean of the k smallest.  Once that bound
// exceeds `reject_at` the hypothesis cannot beat it: the remaining views are
// skipped and kCostInf is returned (callers pass the cost to beat, with a
// rounding margin, or kCostInf to always score fully).  `live` (optional,
// LDS): the bits of a cost that other threads may lower while this one runs,
// each value the full cost of a hypothesis competing with this one; every
// check also rejects against the value it reads, with the same margin.
// Per-hypothesis constants of PlaneCost: the pixel's centre q = (qx, qy, 1)
// and the plane as m = K_r^-T n / rho, with mq = m . q (FP64).  front =
// false: the plane does not face the camera (no view can be valid).
struct HypGeom {
  double qx, qy, D, m0, m1, m2, mq;
  bool front;
};

__device__ __forceinline__ HypGeom MakeHypGeom(const PlaneArgs& a, int x, int y, const Hyp& h) {
  HypGeom g;
  g.qx = x + 0.5;
  g.qy = y + 0.5;
  const double rx = fma(g.qx, a.kinv[0], a.kinv[2]), ry = fma(g.qy, a.kinv[1], a.kinv[3]);
  const double nx = h.nx, ny = h.ny, nz = h.nz;
  g.D = h.d;
  const double nd = nx * rx + ny * ry + nz;
  g.front = nd < 0.0;
  const double irho = 1.0 / (g.D * nd);
  g.m0 = nx * a.kinv[0] * irho;
  g.m1 = ny * a.kinv[1] * irho;
  g.m2 = (nz + nx * a.kinv[2] + ny * a.kinv[3]) * irho;
  g.mq = nd * irho;  // m0 qx + m1 qy + m2
  return g;
}

// 1/x within an ulp of the rounded quotient: the hardware reciprocal
// refined by two Newton steps (x finite, normal and nonzero).
__device__ __forceinline__ double RcpF64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// Costs of the views seen so far: count, sum and the sorted k smallest.
struct CostAcc {
  float topk[kMaxTopKNet];
  float csum;
  int nvalid;
};

__device__ __forceinline__ void ResetCostAcc(CostAcc& acc) {
#pragma unroll
  for (int k = 0; k < kMaxTopKNet; ++k) acc.topk[k] = kCostInf;
  acc.csum = 0.0f;
  acc.nvalid = 0;
}

// Inserts cost c into the sorted k smallest.  Costs are finite and
// non-negative (1 - NCC with NCC clamped to [-1, 1], or kCostInf), so their
// bit patterns order as unsigned integers: integer min/max, which need no
// NaN canonicalisation.  With top_k <= 4 only the first 4 slots are ever
// read (the rest stay kCostInf), so the network stops there.
__device__ __forceinline__ void InsertTopK(float* topk, float c, bool four) {
  uint32_t cu = __float_as_uint(c);
  if (four) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const uint32_t t = __float_as_uint(topk[k]);
      topk[k] = __uint_as_float(min(t, cu));
      cu = max(t, cu);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k) {
      const uint32_t t = __float_as_uint(topk[k]);
      topk[k] = __uint_as_float(min(t, cu));
      cu = max(t, cu);
    }
  }
}

// View v's NCC of the hypothesis (NaN when the view is invalid); a valid
// view's cost 1 - NCC is added to `acc`.
//
// Sample positions are offsets from the window centre's projection (fu, fv)
// (FP64-composed warp terms g, h2 relative to that centre):
//   t = (f, g) + (g0. dx + g0. dy, g1. dx + g1. dy) / (1 + h20 dx + h21 dy).
// A lane is "fast" when its whole window lies inside the source with a
// 1e-4 px margin (a rigorous bound on the warp) and inside the reference: its
// samples need no masks.  When every lane of the wave is fast the window runs
// a loop without masks; otherwise a loop that checks each sample, in which
// fast lanes perform exactly the operations of the mask-free loop -- a
// pixel's result does not depend on which pixels share its wave (tile
// schedules, view-phased sweeps and batches agree bit for bit).
//
// Sums: source values are shifted by the window's first valid sample,
// reference values are those of the LDS tile and their centred sums; every
// sum stays near its own scale, so the FP32 variances and covariance do not
// cancel.
template <int HALF, int STRIDE, int FMT>
__device__ __forceinline__ float ViewNcc(const PlaneArgs& a, int v, const HypGeom& hg,
                                         const RefWindow<HALF, STRIDE>& win, bool use_topk,
                                         CostAcc& acc) {
  constexpr int NS = RefWindow<HALF, STRIDE>::NS;
  constexpr int S = RefWindow<HALF, STRIDE>::S;
  constexpr int kEB = ElemBytes<FMT>();
  constexpr int kLogEB = kEB == 8 ? 3 : kEB == 4 ? 2 : 1;
  const double qx = hg.qx, qy = hg.qy, D = hg.D, m0 = hg.m0, m1 = hg.m1, mq = hg.mq;
  float* topk = acc.topk;
  float& csum = acc.csum;
  int& nvalid = acc.nvalid;
  float ncc_v = __builtin_nanf("");
  // centre projection hc = H q, H = A + b m^T, as A q + b mq: each product
  // and sum takes at most one per-view (uniform) operand, which VALU
  // instructions read from scalar registers without copies
  const double* Av = a.A[v];
  const double* bv = a.b[v];
  const double hc0 = fma(Av[1], qy, fma(Av[0], qx, bv[0] * mq)) + Av[2];
  const double hc1 = fma(Av[4], qy, fma(Av[3], qx, bv[1] * mq)) + Av[5];
  const double hc2 = fma(Av[7], qy, fma(Av[6], qx, bv[2] * mq)) + Av[8];
  if (hc2 * D > 1e-6) {
    const double ih = RcpF64(hc2);
    const double uc = hc0 * ih, vc = hc1 * ih;       // continuous source coords
    const double su = uc - 0.5, sv = vc - 0.5;       // sampler coords (pixel centres at integers)
    if (fabs(su) < 1e7 && fabs(sv) < 1e7) {
      const double fsu = floor(su), fsv = floor(sv);
      const int iu = static_cast<int>(fsu), iv = static_cast<int>(fsv);
      const float fu = static_cast<float>(su - fsu), fv = static_cast<float>(sv - fsv);
      // the first two columns of H
      const double H00 = Av[0] + bv[0] * m0, H01 = Av[1] + bv[0] * m1;
      const double H10 = Av[3] + bv[1] * m0, H11 = Av[4] + bv[1] * m1;
      const double H20 = Av[6] + bv[2] * m0, H21 = Av[7] + bv[2] * m1;
      const float g00 = static_cast<float>(fma(-uc, H20, H00) * ih);
      const float g01 = static_cast<float>(fma(-uc, H21, H01) * ih);
      const float g10 = static_cast<float>(fma(-vc, H20, H10) * ih);
      const float g11 = static_cast<float>(fma(-vc, H21, H11) * ih);
      const float h20 = static_cast<float>(H20 * ih), h21 = static_cast<float>(H21 * ih);
      const int w = a.src_w[v], hgt = a.src_h[v], pitch = a.src_pitch[v];
      const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
          const_cast<void*>(a.src_img[v]), 0, pitch * hgt * kEB, 0x00020000);
      // byte offset of the window-centre element (iv, iu) and of one element
      // row (images stay below 2^24 elements: both fit in 32 bits, the row
      // step in 24)
      const int obase = (iv * pitch + iu) * kEB;
      const int pitch_eb = pitch * kEB;
      // A window entirely beyond one image edge cannot give a valid view:
      // skip it.  Over the window |g0. dx + g0. dy| <= e0 (e1 for v) and
      // the denominator is >= 1 - e2, so |du| <= e0 / (1 - e2).
      const float e0 = HALF * (fabsf(g00) + fabsf(g01)), e1 = HALF * (fabsf(g10) + fabsf(g11));
      const float e2 = HALF * (fabsf(h20) + fabsf(h21));
      const float rd = __builtin_amdgcn_rcpf(1.0f - e2);
      const float bu = e0 * rd + 1e-3f;
      const float bv = e1 * rd + 1e-3f;
      // e2 < 0.5 keeps every sample's denominator positive
      const bool den_ok = e2 < 0.5f;
      // the window's sampler-coordinate box, [su - bu, su + bu] x [sv - bv,
      // sv + bv], against the source (FP64: the limits are per-view uniforms)
      const double xl = su - static_cast<double>(bu), xh = su + static_cast<double>(bu);
      const double yl = sv - static_cast<double>(bv), yh = sv + static_cast<double>(bv);
      const double* lim = a.src_lim[v];
      const bool outside = den_ok & ((xh < 0.0) | (xl > lim[0]) | (yh < 0.0) | (yl > lim[1]));
      const bool fast = den_ok & win.full & (xl >= 1e-4) & (xh <= lim[2]) & (yl >= 1e-4) & (yh <= lim[3]);
      // source sums (ss, sss, srs: values shifted by k_src) and the
      // reference values of excluded samples, centred (xr, xrr), with their
      // count xcnt
      float ss = 0.0f, sss = 0.0f, srs = 0.0f, xr = 0.0f, xrr = 0.0f, xcnt = 0.0f;
      float k_src = 0.0f;
      if (!outside) {
        if (__ballot(!fast) == 0) {
          // every lane is fast: no per-sample masks.  Software-pipelined:
          // row j + 1's texels are requested before row j's are
          // interpolated and summed (the same operations, in the same order,
          // as one row at a time)
          auto row_taps = [&](int j, Texel<FMT>* q, float* fxs, float* fys) {
            const float dy = static_cast<float>(-HALF + j * STRIDE);
            const float b0 = g01 * dy, b1 = g11 * dy, b2 = fmaf(h21, dy, 1.0f);
            int off[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float dx = static_cast<float>(-HALF + i * STRIDE);
              const float r = __builtin_amdgcn_rcpf(fmaf(h20, dx, b2));
              const float tx = fmaf(fmaf(g00, dx, b0), r, fu);
              const float ty = fmaf(fmaf(g10, dx, b1), r, fv);
              fxs[i] = __builtin_amdgcn_fractf(tx);
              fys[i] = __builtin_amdgcn_fractf(ty);
              off[i] = MadI24(FloorInt(ty), pitch_eb, (FloorInt(tx) << kLogEB) + obase);
            }
#pragma unroll
            for (int i = 0; i < NS; ++i) q[i] = LoadTexel<FMT>(rsrc, off[i]);
          };
          auto sum_row = [&](int j, const Texel<FMT>* q, const float* fxs, const float* fys) {
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float b = Bilinear<FMT>(q[i], fxs[i], fys[i]);
              if (i == 0 && j == 0) k_src = b;
              const float s = b - k_src;
              ss += s;
              sss = fmaf(s, s, sss);
              srs = fmaf(win.at(i, j), s, srs);
            }
          };
          // two row buffers used in turn (rows of even j in a, odd j in b)
          Texel<FMT> qa[NS], qb[NS];
          float fxa[NS], fya[NS], fxb[NS], fyb[NS];
          row_taps(0, qa, fxa, fya);
          int j = 0;
#pragma unroll 1
          for (; j + 2 < NS; j += 2) {
            row_taps(j + 1, qb, fxb, fyb);
            sum_row(j, qa, fxa, fya);
            row_taps(j + 2, qa, fxa, fya);
            sum_row(j + 1, qb, fxb, fyb);
          }
          if constexpr (NS % 2 == 0) {
            row_taps(NS - 1, qb, fxb, fyb);
            sum_row(NS - 2, qa, fxa, fya);
            sum_row(NS - 1, qb, fxb, fyb);
          } else {
            sum_row(NS - 1, qa, fxa, fya);
          }
        } else if (__ballot(!(win.full & den_ok)) == 0) {
          // some lane is not fast, but every lane's reference window is
          // whole and every denominator positive: only the source range is
          // checked per sample (relative to (iu, iv)); same operations as
          // the general loop below for such lanes
          const float lox = static_cast<float>(-iu), hix = static_cast<float>(w - 1 - iu);
          const float loy = static_cast<float>(-iv), hiy = static_cast<float>(hgt - 1 - iv);
          // the shift is the first valid sample; NaN until one is seen
          k_src = __builtin_nanf("");
          int ixcnt = 0;
          // software-pipelined as the loop above
          auto row_taps = [&](int j, Texel<FMT>* q, float* txs, float* tys) {
            const float dy = static_cast<float>(-HALF + j * STRIDE);
            const float b0 = g01 * dy, b1 = g11 * dy, b2 = fmaf(h21, dy, 1.0f);
            int off[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float dx = static_cast<float>(-HALF + i * STRIDE);
              const float r = __builtin_amdgcn_rcpf(fmaf(h20, dx, b2));
              txs[i] = fmaf(fmaf(g00, dx, b0), r, fu);
              tys[i] = fmaf(fmaf(g10, dx, b1), r, fv);
              off[i] = MadI24(FloorInt(tys[i]), pitch_eb, (FloorInt(txs[i]) << kLogEB) + obase);
            }
#pragma unroll
            for (int i = 0; i < NS; ++i) q[i] = LoadTexel<FMT>(rsrc, off[i]);
          };
          auto sum_row = [&](int j, const Texel<FMT>* q, const float* txs, const float* tys) {
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float b = Bilinear<FMT>(q[i], __builtin_amdgcn_fractf(txs[i]), __builtin_amdgcn_fractf(tys[i]));
              const float rv = win.at(i, j);
              const bool ok = fast | ((txs[i] >= lox) & (txs[i] <= hix) & (tys[i] >= loy) & (tys[i] <= hiy));
              k_src = (ok & (k_src != k_src)) ? b : k_src;
              const float s = ok ? b - k_src : 0.0f;
              const float x = ok ? 0.0f : rv - win.mean;
              ss += s;
              sss = fmaf(s, s, sss);
              srs = fmaf(rv, s, srs);
              xr += x;
              xrr = fmaf(x, x, xrr);
              ixcnt += ok ? 0 : 1;
            }
          };
          Texel<FMT> qa[NS], qb[NS];
          float txa[NS], tya[NS], txb[NS], tyb[NS];
          row_taps(0, qa, txa, tya);
          int j = 0;
#pragma unroll 1
          for (; j + 2 < NS; j += 2) {
            row_taps(j + 1, qb, txb, tyb);
            sum_row(j, qa, txa, tya);
            row_taps(j + 2, qa, txa, tya);
            sum_row(j + 1, qb, txb, tyb);
          }
          if constexpr (NS % 2 == 0) {
            row_taps(NS - 1, qb, txb, tyb);
            sum_row(NS - 2, qa, txa, tya);
            sum_row(NS - 1, qb, txb, tyb);
          } else {
            sum_row(NS - 1, qa, txa, tya);
          }
          xcnt = static_cast<float>(ixcnt);
        } else {
          // some lane is not fast: each sample is checked; fast lanes take
          // the operations of the loop above
          const float hiy = static_cast<float>(hgt - 1 - iv), loy = static_cast<float>(-iv);
          const float hix = static_cast<float>(w - 1 - iu), lox = static_cast<float>(-iu);
          bool have_k = false;
#pragma unroll 1
          for (int j = 0; j < NS; ++j) {
            const float dy = static_cast<float>(-HALF + j * STRIDE);
            const float b0 = g01 * dy, b1 = g11 * dy, b2 = fmaf(h21, dy, 1.0f);
            float txs[NS], tys[NS], dens[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float dx = static_cast<float>(-HALF + i * STRIDE);
              dens[i] = fmaf(h20, dx, b2);
              const float r = __builtin_amdgcn_rcpf(dens[i]);
              txs[i] = fmaf(fmaf(g00, dx, b0), r, fu);
              tys[i] = fmaf(fmaf(g10, dx, b1), r, fv);
            }
            int off[NS];
            float fxs[NS], fys[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              fxs[i] = __builtin_amdgcn_fractf(txs[i]);
              fys[i] = __builtin_amdgcn_fractf(tys[i]);
              off[i] = MadI24(FloorInt(tys[i]), pitch_eb, (FloorInt(txs[i]) << kLogEB) + obase);
            }
            Texel<FMT> q[NS];
#pragma unroll
            for (int i = 0; i < NS; ++i) q[i] = LoadTexel<FMT>(rsrc, off[i]);
#pragma unroll
            for (int i = 0; i < NS; ++i) {
              const float b = Bilinear<FMT>(q[i], fxs[i], fys[i]);
              const float rv = win.at(i, j);
              const bool rin = rv == rv;  // NaN: outside the reference image
              const bool ok = fast | (rin & (dens[i] > 0.0f) & (txs[i] >= lox) & (txs[i] <= hix) &
                                      (tys[i] >= loy) & (tys[i] <= hiy));
              // the shift: the first valid sample (a fast lane's first sample,
              // as in the loop above)
              if (ok && !have_k) k_src = b;
              have_k = have_k || ok;
              const float s = ok ? b - k_src : 0.0f;
              const float x = (ok || !rin) ? 0.0f : rv - win.mean;
              ss += s;
              sss = fmaf(s, s, sss);
              srs = fmaf(ok ? rv : 0.0f, s, srs);
              xr += x;
              xrr = fmaf(x, x, xrr);
              xcnt += (ok || !rin) ? 0.0f : 1.0f;
            }
          }
        }
      } else {
        xcnt = win.n;
      }
      const float cnt = win.n - xcnt;
      if (cnt >= static_cast<float>(S / 2) && cnt > 0.0f) {
        const float in = __builtin_amdgcn_rcpf(cnt);
        const float mrc = (win.sum_c - xr) * in;       // reference mean - win.mean
        const float ms = ss * in;
        const float vr = fmaf(win.sum_c2 - xrr, in, -mrc * mrc);
        const float vs = fmaf(sss, in, -ms * ms);
        const float cv = fmaf(srs, in, -(mrc + win.mean) * ms);
        const float d2 = vr * vs;
        if (d2 >= a.eps2 && d2 > 0.0f) {
          float ncc = cv * __builtin_amdgcn_rsqf(d2);
          ncc = fminf(fmaxf(ncc, -1.0f), 1.0f);
          ncc_v = ncc;
          float c = 1.0f - ncc;
          ++nvalid;
          csum += c;
          if (use_topk) InsertTopK(topk, c, a.top_k <= 4);
        }
      }
    }
  }
  return ncc_v;
}

// Lower bound of the final cost after view v of a.nsrc, every view still to
// come taken as valid with cost 0 (adding such a view never raises the mean
// of the k smallest).  It is 0 while top_k or more views are still to come
// (then either the k smallest can all be those zeros, or no view has been
// seen), so those views skip it.  The divisions are reciprocal products: their
// rounding (~2e-7 relative; none for a power-of-two top_k) lies far inside the
// callers' margin (the bound is compared with the current cost x (1 + 1e-5)).
// The early-exit threshold for beating cost c: c with a rounding margin
// (kCostInf or above: never exit).
__device__ __forceinline__ float RejectAbove(float c) {
  return c < kCostInf ? fmaf(c, 1.0f + 1e-5f, 1e-7f) : kCostInf;
}

__device__ __forceinline__ float CostLowerBound(const PlaneArgs& a, int v, const CostAcc& acc) {
  const int u = a.nsrc - 1 - v;  // views still to come
  if (u >= a.top_k) return 0.0f;
  float lb = 0.0f;
  if (acc.nvalid + u <= a.top_k) {
    // every valid view is aggregated: the mean of the seen costs and u zeros
    if (acc.nvalid > 0) lb = acc.csum * __builtin_amdgcn_rcpf(static_cast<float>(acc.nvalid + u));
  } else {
    // the k smallest: u zeros and the (k - u) smallest seen costs
    const int need = a.top_k - u;
    float sk = 0.0f;
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k) sk += (k < need) ? acc.topk[k] : 0.0f;
    if (need > 0) lb = sk * __builtin_amdgcn_rcpf(static_cast<float>(a.top_k));
  }
  return lb;
}

// The aggregated cost of all views: mean of the k smallest view costs (or of
// all when top_k covers every view); kCostInf without a valid view or above
// the photometric rejection threshold.
__device__ __forceinline__ float FinishCost(const PlaneArgs& a, const CostAcc& acc, bool use_topk) {
  if (acc.nvalid == 0) return kCostInf;
  float cost;
  if (use_topk) {
    const int kk = min(a.top_k, acc.nvalid);
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k) sum += (k < kk) ? acc.topk[k] : 0.0f;
    cost = sum / static_cast<float>(kk);
  } else {
    cost = acc.csum / static_cast<float>(acc.nvalid);
  }
  return cost > a.photo_reject ? kCostInf : cost;
}

template <int HALF, int STRIDE, int FMT>
__device__ float PlaneCost(const PlaneArgs& a, int x, int y, const RefWindow<HALF, STRIDE>& win,
                           const Hyp& h, int* valid_views_out, float* view_ncc,
                           float reject_at = kCostInf, const int* live = nullptr) {
  *valid_views_out = 0;
  if (!(h.d >= a.min_depth && h.d <= a.max_depth)) {
    if (view_ncc)
      for (int v = 0; v < a.nsrc; ++v) view_ncc[v] = __builtin_nanf("");
    return kCostInf;
  }
  const HypGeom hg = MakeHypGeom(a, x, y, h);
  const bool use_topk = a.top_k < a.nsrc;
  CostAcc acc;
  ResetCostAcc(acc);
  if (hg.front) {
    for (int v = 0; v < a.nsrc; ++v) {
      const float ncc_v = ViewNcc<HALF, STRIDE, FMT>(a, v, hg, win, use_topk, acc);
      if (view_ncc) view_ncc[v] = ncc_v;
      if (live) reject_at = fminf(reject_at, RejectAbove(__int_as_float(__hip_atomic_load(
                                                 live, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))));
      if (reject_at < kCostInf && CostLowerBound(a, v, acc) > reject_at) return kCostInf;
    }
  } else if (view_ncc) {
    for (int v = 0; v < a.nsrc; ++v) view_ncc[v] = __builtin_nanf("");
  }
  *valid_views_out = acc.nvalid;
  return FinishCost(a, acc, use_topk);
}

// ---------------------------------------------------------------------------
// Candidate planes (FP32, no contraction)
// ---------------------------------------------------------------------------

// Ray of pixel (x, y) through its centre, z = 1.
__device__ __forceinline__ void PixelRay(const PlaneArgs& a, int x, int y, float* r) {
  r[0] = (static_cast<float>(x) + 0.5f - a.fcx) / a.ffx;
  r[1] = (static_cast<float>(y) + 0.5f - a.fcy) / a.ffy;
  r[2] = 1.0f;
}

__device__ __forceinline__ bool Normalize3(float* n) {
  const float l2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
  if (!(l2 > 1e-12f)) return false;
  const float l = sqrtf(l2);
  n[0] = n[0] / l;
  n[1] = n[1] / l;
  n[2] = n[2] / l;
  return true;
}

// Random plane: inverse depth uniform in [inv_min, inv_max]; normal from a
// uniform cube sample, normalised and turned towards the camera.
__device__ Hyp RandomPlane(const PlaneArgs& a, uint32_t key, uint32_t k0, const float* ray) {
  const float u0 = RngUniform(key, k0);
  const float inv = a.inv_min + u0 * (a.inv_max - a.inv_min);
  float n[3] = {2.0f * RngUniform(key, k0 + 1) - 1.0f, 2.0f * RngUniform(key, k0 + 2) - 1.0f,
                2.0f * RngUniform(key, k0 + 3) - 1.0f};
  if (!Normalize3(n)) {
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
  return Hyp{1.0f / inv, n[0], n[1], n[2]};
}

}  // namespace rmvs

using namespace rmvs;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Tile index -> (tile x, tile y), XCD-aware: workgroups are dealt round-robin
// to the 8 XCDs, so consecutive tile indices are placed on one XCD (shared
// L2 for their overlapping source footprints), and tile indices run down the
// columns: each XCD holds a vertical band of the image, and the tiles in
// flight on it at one time form a roughly square block (about 240 x 200 px at
// 1920 x 1440).  A hypothesis at an unknown depth samples a source along its
// epipolar line; a square block keeps the union of those segments compact
// for every baseline direction, where a band of rows does so only for
// sources displaced along the rows.
__device__ __forceinline__ bool TileOf(const PlaneArgs& a, int* tx, int* ty) {
  const int b = blockIdx.x;
  const int t = (b & 7) * a.tiles_per_xcd + (b >> 3);
  if (t >= a.num_tiles) return false;
  *tx = t / a.tiles_y;
  *ty = t % a.tiles_y;
  return true;
}

// Reference tile of a workgroup: TW x TH pixels plus a HALF-pixel halo in
// LDS, every value shifted by the value at the centre of the 128 x 16 block
// that holds the tile (NaN outside the image).  256 threads; a thread owns
// one pixel of each colour: thread t -> row t / (TW/2), the (t % (TW/2))-th
// pixel of the colour in that row.
template <int HALF, int FMT, int TW, int TH>
struct RefTile {
  static_assert(TW * TH == 512, "256 threads, two pixels each");
  static_assert(128 % TW == 0 && 16 % TH == 0, "a tile lies inside one 128 x 16 block");
  static constexpr int LW = TW + 2 * HALF + 1;  // +1: odd row pitch against bank conflicts
  static constexpr int LH = TH + 2 * HALF;
  float v[LH][LW];
  float shift;

  __device__ __forceinline__ void load(const PlaneArgs& a, int x0, int y0) {
    if (threadIdx.x == 0) {
      // the centre of the 128 x 16 block holding the tile: every tile
      // schedule (64x8, 32x16, 128x4) lies inside one such block, so a pixel
      // sees the same shift, and the same rounding, under every schedule
      const int cx = min((x0 & ~127) + 64, a.w - 1), cy = min((y0 & ~15) + 8, a.h - 1);
      shift = Pix<FMT>(a.ref_img, cy * a.ref_pitch + cx);
    }
    __syncthreads();
    const float sh = shift;
    for (int i = threadIdx.x; i < (LW - 1) * LH; i += blockDim.x) {
      const int ly = i / (LW - 1), lx = i % (LW - 1);
      const int gx = x0 - HALF + lx, gy = y0 - HALF + ly;
      float val = __builtin_nanf("");
      if (gx >= 0 && gy >= 0 && gx < a.w && gy < a.h) val = Pix<FMT>(a.ref_img, gy * a.ref_pitch + gx) - sh;
      v[ly][lx] = val;
    }
  }
  // pixel of colour `color` owned by thread t
  __device__ __forceinline__ void pixel_of(int t, int x0, int y0, int color, int* x, int* y, int* r) const {
    *r = t / (TW / 2);
    const int j = t % (TW / 2);
    *y = y0 + *r;
    *x = x0 + 2 * j + (((x0 + *y) & 1) ^ color);
  }
  // pixel of colour `color` owned by this thread
  __device__ __forceinline__ void pixel(int x0, int y0, int color, int* x, int* y, int* r) const {
    pixel_of(static_cast<int>(threadIdx.x), x0, y0, color, x, y, r);
  }
  template <int STRIDE>
  __device__ __forceinline__ void window(RefWindow<HALF, STRIDE>& win, int x, int x0, int r) const {
    MakeRefWindow<HALF, STRIDE>(win, &v[r][x - x0], STRIDE * LW, shift);
  }
};

// ---------------------------------------------------------------------------
// Host helpers (defined in plane_mode.hip)
// ---------------------------------------------------------------------------
namespace rmvs {

// Reference-to-source rigid motions of one problem: Y = R_rel X + t_rel.
struct PairGeometry {
  double Rrel[kMaxViews][9];
  double trel[kMaxViews][3];
};

// Rotation matrix Exp([w]x) of a rotation vector (Rodrigues, FP64).
void So3Exp(const double w[3], double E[9]);

// Kernel arguments for reference view `ref` and its sources.  pose_delta
// (nsrc x 6, may be null) perturbs each source's world->camera transform on
// the left, Y -> Exp(w) Y + tau with (w, tau) = pose_delta[6k .. 6k+5];
// geo (may be null) receives the resulting R_rel, t_rel.
int32_t PlaneMakeArgs(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs,
                      const recon_mvs_plane_params* p, PlaneArgs* a,
                      const double* pose_delta = nullptr, PairGeometry* geo = nullptr);
// Allocates the depth-map slot of `view` (no-op when it exists with its size).
int32_t PlaneEnsureState(recon_mvs_hip_context* ctx, int view);
// Mean reference-to-source baseline (patchmatch_stereo.cpp:472-479).
double PlaneMeanBaseline(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs);

}  // namespace rmvs
