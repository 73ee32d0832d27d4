// plane_adjoint.hip — reverse mode of the plane-mode cost on MI355X.
//
// For every pixel of a depth-map slot the aggregated cost C_p (mean of 1 - NCC
// over the selected views, as in PlaneCost) is differentiated with respect to
// the pixel's plane (depth D, normal n) and the six pose parameters of every
// source view.  A source pose is perturbed on the left, Y -> Exp(w) Y + tau
// for Y in source-camera coordinates, so dY/dw = -[Y]x and dY/dtau = I.
//
// Per sample i of view v (window offset (dx, dy), ray r_i, plane point
// X_i = lambda_i r_i with lambda_i = D (n.r_c) / (n.r_i), Y_i = R_rel X_i +
// t_rel, source pixel u_i = f Y / Y_z + c):
//   d NCC / d s_i = (r~_i / (sigma_r sigma_s) - NCC s~_i / sigma_s^2) / N
//   d s / d(u, v)  = bilinear slopes of the four taps
//   Y_z g_Y        = (g_u f_x, g_v f_y, -(g_u (u - c_x) + g_v (v - c_y)))
//   dC/dw += Y x g_Y,  dC/dtau += g_Y,  g_X = R_rel^T g_Y
//   dC/dD += (g_X . r_i) lambda_i / D
//   dC/dn += (g_X . r_i) / (n . r_i) (D r_c - lambda_i r_i)
// The warp is recomputed from the plane (no stored warps): one pass per view
// for the statistics that pick the top-k views, then for each selected view
// one statistics pass and one gradient pass.  Per-pixel sums are FP32; the
// pose gradient is reduced in FP64 (wave shuffles, then per-wave LDS slots,
// then per-tile partials summed in tile order), so it is deterministic.
//
// The same machinery drives a gradient polish of a finished depth map: each
// step moves every pixel's inverse depth (even steps) or normal (odd steps)
// against its gradient by a per-pixel step length, keeps the move only if the
// cost drops, and grows or halves that step length.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mvs_common.hpp"
#include "plane_common.hpp"

namespace {

constexpr int kAdjTW = 32, kAdjTH = 16;  // tile of the adjoint kernels (256 threads)

// Scratch of recon_mvs_hip_plane_cost_grad_noalloc: the per-tile FP64 pose
// partials (tiles x num_src x 6 doubles) in a fixed 2 MiB device array -- a
// 1920x1440 map (5400 tiles of 32 x 16) against 8 sources.  Each GPU the
// library runs on holds its own copy until the process ends.  Host code uses
// the address AdjointScratchInit obtained for the device, never this name.
constexpr size_t kNoAllocScratchBytes = size_t{2} << 20;
static __device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];
constexpr int kTimerAdjoint = 7;

struct AdjArgs {
  PlaneArgs a;
  float rrel[kMaxViews][9];  // R_rel, FP32
  float ks[kMaxViews][4];    // source fx, fy, cx, cy
  const float* weights;      // per-pixel weight of the summed cost, null = 1
  float* cost_out;           // per-pixel cost (kCostInf: invalid), may be null
  float4* grad_plane;        // per-pixel (dD, dn), may be null
  double* pose_part;         // per tile [nsrc][6], may be null
  float2* steps;             // polish: per-pixel step lengths (inverse depth, normal)
  float inv_per_px;          // polish: inverse depth of one pixel of disparity
  int step;                  // polish: step index
};

// Per-view FP32 coefficients of one plane at one pixel (FP64 composed).
struct ViewGeom {
  float fu, fv;                      // fractional part of the centre's sampler coordinates
  float g00, g01, g10, g11, h20, h21;
  float lox, hix, loy, hiy;
  float obase, fpitch;
  float ucx, vcy;                    // centre projection minus the source principal point
  float inv_rh;                      // 1 / (rho hc2): Y_z = lambda hc2 den
  float qc[3];                       // R_rel r_c
};

struct ViewStats {
  float n, mr, ms, var_r, var_s, ncc, raw;
};

// FP64 set-up of view v for plane (D, n) at pixel (qx, qy) = (x + .5, y + .5)
// (the same composition as PlaneCost).  False: the view is rejected.
__device__ __forceinline__ bool SetupView(const PlaneArgs& a, const AdjArgs& g, int v, double qx,
                                          double qy, double D, double m0, double m1, double m2,
                                          double rho, const double rc[3], int HALF_, ViewGeom* vg) {
  double H[9];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    H[3 * i + 0] = fma(a.b[v][i], m0, a.A[v][3 * i + 0]);
    H[3 * i + 1] = fma(a.b[v][i], m1, a.A[v][3 * i + 1]);
    H[3 * i + 2] = fma(a.b[v][i], m2, a.A[v][3 * i + 2]);
  }
  const double hc0 = fma(H[0], qx, fma(H[1], qy, H[2]));
  const double hc1 = fma(H[3], qx, fma(H[4], qy, H[5]));
  const double hc2 = fma(H[6], qx, fma(H[7], qy, H[8]));
  if (!(hc2 * D > 1e-6)) return false;
  const double ih = 1.0 / hc2;
  const double uc = hc0 * ih, vc = hc1 * ih;
  const double su = uc - 0.5, sv = vc - 0.5;
  if (!(fabs(su) < 1e7 && fabs(sv) < 1e7)) return false;
  const double fsu = floor(su), fsv = floor(sv);
  const int iu = static_cast<int>(fsu), iv = static_cast<int>(fsv);
  vg->fu = static_cast<float>(su - fsu);
  vg->fv = static_cast<float>(sv - fsv);
  vg->g00 = static_cast<float>((H[0] - uc * H[6]) * ih);
  vg->g01 = static_cast<float>((H[1] - uc * H[7]) * ih);
  vg->g10 = static_cast<float>((H[3] - vc * H[6]) * ih);
  vg->g11 = static_cast<float>((H[4] - vc * H[7]) * ih);
  vg->h20 = static_cast<float>(H[6] * ih);
  vg->h21 = static_cast<float>(H[7] * ih);
  const int w = a.src_w[v], hgt = a.src_h[v];
  vg->lox = static_cast<float>(-iu);
  vg->hix = static_cast<float>(w - 1 - iu);
  vg->loy = static_cast<float>(-iv);
  vg->hiy = static_cast<float>(hgt - 1 - iv);
  vg->fpitch = static_cast<float>(a.src_pitch[v]);
  vg->obase = static_cast<float>(iv) * vg->fpitch + static_cast<float>(iu);
  vg->ucx = static_cast<float>(uc - static_cast<double>(g.ks[v][2]));
  vg->vcy = static_cast<float>(vc - static_cast<double>(g.ks[v][3]));
  vg->inv_rh = static_cast<float>(1.0 / (rho * hc2));
#pragma unroll
  for (int i = 0; i < 3; ++i)
    vg->qc[i] = static_cast<float>(g.rrel[v][3 * i] * rc[0] + g.rrel[v][3 * i + 1] * rc[1] +
                                   static_cast<double>(g.rrel[v][3 * i + 2]));
  (void)HALF_;
  return true;
}

// Gradient sums of one view at one pixel.
struct ViewGrad {
  float A1, A2, B, Bx, By;  // sum a lambda, sum a / nr, sum b, sum b dx, sum b dy
  float rw[3], tw[3];       // pose: sum Y x g_Y, sum g_Y
};

// One pass over the window of view v.  MODE 0: statistics; MODE 1: gradient
// sums with the statistics `st` and d(cost)/d(ncc) = dc.
template <int HALF, int STRIDE, int FMT, int MODE>
__device__ __forceinline__ void ViewPass(const PlaneArgs& a, const AdjArgs& g, int v,
                                         const ViewGeom& vg, const RefWindow<HALF, STRIDE>& win,
                                         float nd, float nxf, float nyf, float rho, ViewStats* st,
                                         float dc, ViewGrad* gr, float ifx, float ify) {
  constexpr int NS = RefWindow<HALF, STRIDE>::NS;
  constexpr int kEB = ElemBytes<FMT>();
  const int pitch = a.src_pitch[v], hgt = a.src_h[v];
  const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<void*>(a.src_img[v]), 0, pitch * hgt * kEB, 0x00020000);
  const float shift = win.shift;
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f, s4 = 0.0f, cnt = 0.0f;
  float al = 0.0f, be = 0.0f;
  if constexpr (MODE == 1) {
    al = __builtin_amdgcn_rcpf(st->n * sqrtf(st->var_r * st->var_s));
    be = st->ncc / (st->n * st->var_s);
  }
  const float fx_s = g.ks[v][0], fy_s = g.ks[v][1];
#pragma unroll 1
  for (int j = 0; j < NS; ++j) {
    const float dy = static_cast<float>(-HALF + j * STRIDE);
    const float b0 = vg.g01 * dy, b1 = vg.g11 * dy, b2 = fmaf(vg.h21, dy, 1.0f);
#pragma unroll
    for (int i = 0; i < NS; ++i) {
      const float dx = static_cast<float>(-HALF + i * STRIDE);
      const float den = fmaf(vg.h20, dx, b2);
      const float rd = 1.0f / den;
      const float du = fmaf(vg.g00, dx, b0) * rd, dv = fmaf(vg.g10, dx, b1) * rd;
      const float tx = vg.fu + du, ty = vg.fv + dv;
      const float rv = win.at(i, j);
      const bool ok = (rv == rv) & (den > 0.0f) & (tx >= vg.lox) & (tx <= vg.hix) &
                      (ty >= vg.loy) & (ty <= vg.hiy);
      const float flx = floorf(tx), fly = floorf(ty);
      const float fx = tx - flx, fy = ty - fly;
      const int off = static_cast<int>(fmaf(fly, vg.fpitch, flx + vg.obase)) * kEB;
      const float4 qd = BufQuad<FMT>(rsrc, ok ? off : 0);
      const float dtop = qd.z - qd.x, dbot = qd.w - qd.y;
      const float top = fmaf(fx, dtop, qd.x), bot = fmaf(fx, dbot, qd.y);
      const float sv = fmaf(fy, bot - top, top) - shift;
      if constexpr (MODE == 0) {
        const float s = ok ? sv : 0.0f, r = ok ? rv : 0.0f;
        s0 += s;
        s1 = fmaf(s, s, s1);
        s2 = fmaf(r, s, s2);
        s3 += r;
        s4 = fmaf(r, r, s4);
        cnt += ok ? 1.0f : 0.0f;
      } else {
        if (ok) {
          const float gs = dc * ((rv - st->mr) * al - (sv - st->ms) * be);
          const float gu = gs * fmaf(fy, dbot - dtop, dtop);
          const float gv = gs * (bot - top);
          const float ub = vg.ucx + du, vb = vg.vcy + dv;  // u - c_x, v - c_y of the sample
          // Y_z g_Y and Y / Y_z
          const float G0 = gu * fx_s, G1 = gv * fy_s, G2 = -(gu * ub + gv * vb);
          const float y0 = ub / fx_s, y1 = vb / fy_s;
          gr->rw[0] += y1 * G2 - G1;
          gr->rw[1] += G0 - y0 * G2;
          gr->rw[2] += y0 * G1 - y1 * G0;
          const float nr = fmaf(nxf, dx, fmaf(nyf, dy, nd));
          const float inr = 1.0f / nr;
          const float iyz = nr * rd * vg.inv_rh;  // 1 / Y_z
          gr->tw[0] += G0 * iyz;
          gr->tw[1] += G1 * iyz;
          gr->tw[2] += G2 * iyz;
          // (R_rel r_i) . g_Y with r_i = r_c + (dx / fx, dy / fy, 0)
          const float q0 = fmaf(g.rrel[v][0], dx * ifx, fmaf(g.rrel[v][1], dy * ify, vg.qc[0]));
          const float q1 = fmaf(g.rrel[v][3], dx * ifx, fmaf(g.rrel[v][4], dy * ify, vg.qc[1]));
          const float q2 = fmaf(g.rrel[v][6], dx * ifx, fmaf(g.rrel[v][7], dy * ify, vg.qc[2]));
          const float ai = (q0 * G0 + q1 * G1 + q2 * G2) * iyz;
          const float lam = rho * inr;
          const float al_i = ai * lam;
          gr->A1 += al_i;
          gr->A2 += ai * inr;
          const float bi = al_i * inr;
          gr->B += bi;
          gr->Bx += bi * dx;
          gr->By += bi * dy;
        }
      }
    }
  }
  if constexpr (MODE == 0) {
    st->n = cnt;
    st->raw = 0.0f;
    st->ncc = 0.0f;
    if (cnt >= static_cast<float>(RefWindow<HALF, STRIDE>::S / 2) && cnt > 0.0f) {
      const float in = 1.0f / cnt;
      const float mr = s3 * in, ms = s0 * in;
      const float vr = fmaf(s4, in, -mr * mr);
      const float vs = fmaf(s1, in, -ms * ms);
      const float cv = fmaf(s2, in, -mr * ms);
      const float d2 = vr * vs;
      st->mr = mr;
      st->ms = ms;
      st->var_r = vr;
      st->var_s = vs;
      if (d2 >= a.eps2 && d2 > 0.0f) {
        st->raw = cv / sqrtf(d2);
        st->ncc = fminf(fmaxf(st->raw, -1.0f), 1.0f);
        return;
      }
    }
    st->n = 0.0f;  // invalid
  }
}

// Plane geometry of pixel (x, y): FP64 m = K_r^-T n / rho and the centre ray.
struct PixelGeom {
  double qx, qy, D, m0, m1, m2, rho, rc[3];
  float nd, nxf, nyf;
  bool ok;
};

__device__ __forceinline__ PixelGeom MakePixelGeom(const PlaneArgs& a, int x, int y, float4 p) {
  PixelGeom pg;
  pg.qx = x + 0.5;
  pg.qy = y + 0.5;
  pg.rc[0] = (pg.qx - a.cx) / a.fx;
  pg.rc[1] = (pg.qy - a.cy) / a.fy;
  pg.rc[2] = 1.0;
  const double nx = p.y, ny = p.z, nz = p.w;
  pg.D = p.x;
  const double nd = nx * pg.rc[0] + ny * pg.rc[1] + nz;
  pg.ok = (p.x >= a.min_depth && p.x <= a.max_depth) && nd < 0.0;
  pg.rho = pg.D * nd;
  pg.m0 = nx / a.fx / pg.rho;
  pg.m1 = ny / a.fy / pg.rho;
  pg.m2 = (nz - nx * a.cx / a.fx - ny * a.cy / a.fy) / pg.rho;
  pg.nd = static_cast<float>(nd);
  pg.nxf = static_cast<float>(nx / a.fx);
  pg.nyf = static_cast<float>(ny / a.fy);
  return pg;
}

// Cost and (dD, dn) of one pixel; per selected view the pose sums are handed
// to `pose(v, grad)` (called by every lane for every view, uniformly, with
// zeros where the lane has nothing).
template <int HALF, int STRIDE, int FMT, bool kPose, typename PoseSink>
__device__ __forceinline__ float PixelCostGrad(const PlaneArgs& a, const AdjArgs& g, bool active,
                                               int x, int y, float4 p, float w,
                                               const RefWindow<HALF, STRIDE>& win, float4* grad,
                                               PoseSink&& pose) {
  *grad = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  PixelGeom pg;
  pg.ok = false;
  if (active) pg = MakePixelGeom(a, x, y, p);
  const bool use_topk = a.top_k < a.nsrc;
  const float ifx = static_cast<float>(1.0 / a.fx), ify = static_cast<float>(1.0 / a.fy);
  float tk[kMaxTopKNet];
  int ti[kMaxTopKNet];
#pragma unroll
  for (int k = 0; k < kMaxTopKNet; ++k) {
    tk[k] = kCostInf;
    ti[k] = 0;
  }
  unsigned valid_mask = 0u;
  float csum = 0.0f;
  int nvalid = 0;
  if (pg.ok) {
    for (int v = 0; v < a.nsrc; ++v) {
      ViewGeom vg;
      if (!SetupView(a, g, v, pg.qx, pg.qy, pg.D, pg.m0, pg.m1, pg.m2, pg.rho, pg.rc, HALF, &vg)) continue;
      ViewStats st;
      ViewPass<HALF, STRIDE, FMT, 0>(a, g, v, vg, win, pg.nd, pg.nxf, pg.nyf,
                                     static_cast<float>(pg.rho), &st, 0.0f, nullptr, ifx, ify);
      if (!(st.n > 0.0f)) continue;
      float c = 1.0f - st.ncc;
      int id = v;
      valid_mask |= 1u << v;
      csum += c;
      ++nvalid;
      // insertion into the sorted k smallest; equal costs keep the lower view first
#pragma unroll
      for (int k = 0; k < kMaxTopKNet; ++k) {
        const bool lt = c < tk[k];
        const float tc = tk[k];
        const int tid = ti[k];
        tk[k] = lt ? c : tc;
        ti[k] = lt ? id : tid;
        c = lt ? tc : c;
        id = lt ? tid : id;
      }
    }
  }
  float cost = kCostInf;
  unsigned sel = 0u;
  int kk = 0;
  if (nvalid > 0) {
    if (use_topk) {
      kk = min(a.top_k, nvalid);
      float sum = 0.0f;
#pragma unroll
      for (int k = 0; k < kMaxTopKNet; ++k) {
        if (k < kk) {
          sum += tk[k];
          sel |= 1u << ti[k];
        }
      }
      cost = sum / static_cast<float>(kk);
    } else {
      kk = nvalid;
      sel = valid_mask;
      cost = csum / static_cast<float>(nvalid);
    }
    if (cost > a.photo_reject) {
      cost = kCostInf;
      sel = 0u;
    }
  }
  const float dc = -w / static_cast<float>(max(kk, 1));
  float gD = 0.0f, gn0 = 0.0f, gn1 = 0.0f, gn2 = 0.0f;
  for (int v = 0; v < a.nsrc; ++v) {
    ViewGrad gr = {};
    if ((sel >> v) & 1u) {
      ViewGeom vg;
      if (SetupView(a, g, v, pg.qx, pg.qy, pg.D, pg.m0, pg.m1, pg.m2, pg.rho, pg.rc, HALF, &vg)) {
        ViewStats st;
        const float rhof = static_cast<float>(pg.rho);
        ViewPass<HALF, STRIDE, FMT, 0>(a, g, v, vg, win, pg.nd, pg.nxf, pg.nyf, rhof, &st, 0.0f, nullptr, ifx, ify);
        // a clamped NCC is flat
        if (st.n > 0.0f && st.raw >= -1.0f && st.raw <= 1.0f) {
          ViewPass<HALF, STRIDE, FMT, 1>(a, g, v, vg, win, pg.nd, pg.nxf, pg.nyf, rhof, &st, dc, &gr, ifx, ify);
          const float D = static_cast<float>(pg.D);
          const float rc0 = static_cast<float>(pg.rc[0]), rc1 = static_cast<float>(pg.rc[1]);
          gD += gr.A1 / D;
          gn0 += D * gr.A2 * rc0 - (gr.B * rc0 + gr.Bx * ifx);
          gn1 += D * gr.A2 * rc1 - (gr.B * rc1 + gr.By * ify);
          gn2 += D * gr.A2 - gr.B;
        }
      }
    }
    if constexpr (kPose) pose(v, gr);
  }
  *grad = make_float4(gD, gn0, gn1, gn2);
  return cost;
}

__device__ __forceinline__ double WaveSum(double x) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  return x;
}

// Cost and gradients of every pixel of one tile.  Grid: XCD-aware tiles of
// 32 x 16 pixels.
template <int HALF, int STRIDE, int FMT>
__global__ __launch_bounds__(256) void k_plane_adjoint(AdjArgs g) {
  const PlaneArgs& a = g.a;
  __shared__ RefTile<HALF, FMT, kAdjTW, kAdjTH> tile;
  __shared__ double spose[4][kMaxViews][6];
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  const int x0 = tx * kAdjTW, y0 = ty * kAdjTH;
  tile.load(a, x0, y0);
  for (int i = threadIdx.x; i < 4 * kMaxViews * 6; i += blockDim.x) (&spose[0][0][0])[i] = 0.0;
  __syncthreads();
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const bool want_pose = g.pose_part != nullptr;
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    tile.pixel(x0, y0, color, &x, &y, &r);
    const bool active = x < a.w && y < a.h;
    const int idx = active ? y * a.w + x : 0;
    RefWindow<HALF, STRIDE> win;
    tile.template window<STRIDE>(win, active ? x : x0, x0, r);
    const float4 p = active ? a.plane[idx] : make_float4(0.0f, 0.0f, 0.0f, -1.0f);
    const float w = (active && g.weights) ? g.weights[idx] : 1.0f;
    float4 grad;
    auto sink = [&](int v, const ViewGrad& gr) {
      if (!want_pose) return;
      const float vals[6] = {gr.rw[0], gr.rw[1], gr.rw[2], gr.tw[0], gr.tw[1], gr.tw[2]};
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        const double s = WaveSum(static_cast<double>(vals[k]));
        if (lane == 0) spose[wave][v][k] += s;
      }
    };
    const float c = PixelCostGrad<HALF, STRIDE, FMT, true>(a, g, active, x, y, p, w, win, &grad, sink);
    if (active) {
      if (g.cost_out) g.cost_out[idx] = c;
      if (g.grad_plane) g.grad_plane[idx] = grad;
    }
  }
  __syncthreads();
  if (want_pose) {
    const int t = ty * a.tiles_x + tx;
    for (int i = threadIdx.x; i < a.nsrc * 6; i += blockDim.x) {
      const int v = i / 6, k = i % 6;
      g.pose_part[static_cast<size_t>(t) * a.nsrc * 6 + i] =
          ((spose[0][v][k] + spose[1][v][k]) + spose[2][v][k]) + spose[3][v][k];
    }
  }
}

// Sums the per-tile pose partials in tile order: one thread per parameter.
__global__ void k_pose_reduce(const double* __restrict__ part, int num_tiles, int n,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int t = 0; t < num_tiles; ++t) s += part[static_cast<size_t>(t) * n + i];
  out[i] = s;
}

// One polish step over every pixel of one tile (see the file comment).
template <int HALF, int STRIDE, int FMT>
__global__ __launch_bounds__(256) void k_plane_polish(AdjArgs g) {
  const PlaneArgs& a = g.a;
  __shared__ RefTile<HALF, FMT, kAdjTW, kAdjTH> tile;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  const int x0 = tx * kAdjTW, y0 = ty * kAdjTH;
  tile.load(a, x0, y0);
  __syncthreads();
  const bool depth_step = (g.step & 1) == 0;
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    tile.pixel(x0, y0, color, &x, &y, &r);
    if (x >= a.w || y >= a.h) continue;
    const int idx = y * a.w + x;
    const float c0 = a.cost[idx];
    float2 st = g.steps[idx];
    if (g.step < 2) {  // first use of this step length
      if (depth_step) st.x = 0.5f * g.inv_per_px;
      else st.y = 0.05f;
    }
    if (c0 < kCostInf) {
      RefWindow<HALF, STRIDE> win;
      tile.template window<STRIDE>(win, x, x0, r);
      const float4 p = a.plane[idx];
      float4 grad;
      auto none = [](int, const ViewGrad&) {};
      const float c = PixelCostGrad<HALF, STRIDE, FMT, false>(a, g, true, x, y, p, 1.0f, win, &grad, none);
      float4 q = p;
      bool moved = false;
      if (c < kCostInf) {
        if (depth_step) {
          // d C / d(1/D) = -D^2 dC/dD: move 1/D by one step length against it
          const float gs = -p.x * p.x * grad.x;
          if (gs != 0.0f) {
            const float inv = 1.0f / p.x - copysignf(st.x, gs);
            if (inv >= a.inv_min && inv <= a.inv_max) {
              q.x = 1.0f / inv;
              moved = true;
            }
          }
        } else {
          // tangential part of dC/dn, then a rotation of the normal by st.y
          const float gdn = grad.y * p.y + grad.z * p.z + grad.w * p.w;
          const float t0 = grad.y - gdn * p.y, t1 = grad.z - gdn * p.z, t2 = grad.w - gdn * p.w;
          const float tn = sqrtf(t0 * t0 + t1 * t1 + t2 * t2);
          if (tn > 0.0f) {
            float n0 = p.y - st.y * t0 / tn, n1 = p.z - st.y * t1 / tn, n2 = p.w - st.y * t2 / tn;
            const float l = sqrtf(n0 * n0 + n1 * n1 + n2 * n2);
            n0 /= l;
            n1 /= l;
            n2 /= l;
            const float ray0 = (static_cast<float>(x) + 0.5f - a.fcx) / a.ffx;
            const float ray1 = (static_cast<float>(y) + 0.5f - a.fcy) / a.ffy;
            if (n0 * ray0 + n1 * ray1 + n2 < 0.0f) {
              q.y = n0;
              q.z = n1;
              q.w = n2;
              moved = true;
            }
          }
        }
      }
      bool accepted = false;
      if (moved) {
        int nv = 0;
        const float cq = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, Hyp{q.x, q.y, q.z, q.w}, &nv, nullptr);
        if (cq < c0) {
          a.plane[idx] = q;
          a.cost[idx] = cq;
          a.support[idx] = static_cast<uint8_t>(min(nv, 255));
          accepted = true;
        }
      }
      if (depth_step) st.x *= accepted ? 1.25f : 0.5f;
      else st.y = accepted ? fminf(st.y * 1.25f, 0.2f) : st.y * 0.5f;
    }
    g.steps[idx] = st;
  }
}

enum AdjKernel { kAdjGrad = 0, kAdjPolish = 1 };

template <int HALF, int STRIDE, int FMT>
void LaunchAdjT(AdjKernel k, const AdjArgs& g, dim3 grid, hipStream_t s) {
  if (k == kAdjGrad) hipLaunchKernelGGL((k_plane_adjoint<HALF, STRIDE, FMT>), grid, dim3(256), 0, s, g);
  else hipLaunchKernelGGL((k_plane_polish<HALF, STRIDE, FMT>), grid, dim3(256), 0, s, g);
}

// Window shapes of the adjoint: (11, 2) in every storage format, (7, 1) and
// (5, 1) in FP32.
int32_t LaunchAdj(recon_mvs_hip_context* ctx, const recon_mvs_plane_params* p, AdjKernel k,
                  AdjArgs* g, bool timed = true) {
  PlaneArgs& a = g->a;
  a.tiles_x = (a.w + kAdjTW - 1) / kAdjTW;
  a.tiles_y = (a.h + kAdjTH - 1) / kAdjTH;
  a.num_tiles = a.tiles_x * a.tiles_y;
  a.tiles_per_xcd = (a.num_tiles + 7) / 8;
  const dim3 grid(a.tiles_per_xcd * 8);
  const int fmt = ctx->storage_format, ps = p->patch_size, st = p->window_stride;
  hipEvent_t eb = nullptr, ee = nullptr;
  if (timed) rmvs::TimerBegin(ctx, kTimerAdjoint, &eb, &ee);
  bool ok = true;
  if (ps == 11 && st == 2) {
    if (fmt == RECON_MVS_PIXEL_F32) LaunchAdjT<5, 2, RECON_MVS_PIXEL_F32>(k, *g, grid, ctx->stream);
    else if (fmt == RECON_MVS_PIXEL_F16) LaunchAdjT<5, 2, RECON_MVS_PIXEL_F16>(k, *g, grid, ctx->stream);
    else if (fmt == RECON_MVS_PIXEL_U8) LaunchAdjT<5, 2, RECON_MVS_PIXEL_U8>(k, *g, grid, ctx->stream);
    else LaunchAdjT<5, 2, RECON_MVS_PIXEL_BF16>(k, *g, grid, ctx->stream);
  } else if (ps == 7 && st == 1 && fmt == RECON_MVS_PIXEL_F32) {
    LaunchAdjT<3, 1, RECON_MVS_PIXEL_F32>(k, *g, grid, ctx->stream);
  } else if (ps == 5 && st == 1 && fmt == RECON_MVS_PIXEL_F32) {
    LaunchAdjT<2, 1, RECON_MVS_PIXEL_F32>(k, *g, grid, ctx->stream);
  } else {
    ok = false;
  }
  if (!ok)
    return rmvs::Fail(ctx, RECON_MVS_ERR_NOT_IMPLEMENTED,
                      "adjoint: window (patch_size, window_stride) / storage format not compiled in");
  if (timed) rmvs::TimerEnd(ctx, kTimerAdjoint, eb, ee);
  RMVS_LAUNCH_CHECK(ctx);
  return RECON_MVS_OK;
}

// Device address of `scratch` per device, fetched once per device by
// AdjointScratchInit (with that device current).
constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxScratchDevices] = {};

int32_t MakeAdjArgs(recon_mvs_hip_context* ctx, int32_t ref, int32_t nsrc, const int32_t* srcs,
                    const recon_mvs_plane_params* p, const double* pose_delta, AdjArgs* g) {
  std::memset(g, 0, sizeof(*g));
  PairGeometry geo;
  int32_t st = rmvs::PlaneMakeArgs(ctx, ref, nsrc, srcs, p, &g->a, pose_delta, &geo);
  if (st != RECON_MVS_OK) return st;
  if (ref >= static_cast<int>(ctx->states.size()) || !ctx->states[ref])
    return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "no depth map in this slot");
  const PlaneState& S = *ctx->states[ref];
  g->a.plane = S.plane;
  g->a.cost = S.cost;
  g->a.support = S.support;
  for (int k = 0; k < nsrc; ++k) {
    for (int i = 0; i < 9; ++i) g->rrel[k][i] = static_cast<float>(geo.Rrel[k][i]);
    const recon_mvs_camera& c = ctx->views[srcs[k]]->cam;
    g->ks[k][0] = static_cast<float>(c.fx);
    g->ks[k][1] = static_cast<float>(c.fy);
    g->ks[k][2] = static_cast<float>(c.cx);
    g->ks[k][3] = static_cast<float>(c.cy);
  }
  return RECON_MVS_OK;
}

}  // namespace

namespace rmvs {

int32_t AdjointScratchInit(int device) {
  if (device < 0 || device >= kMaxScratchDevices) return RECON_MVS_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return RECON_MVS_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
    return RECON_MVS_ERR_GPU_FAILURE;
  g_scratch_ptr[device] = ptr;
  return RECON_MVS_OK;
}

}  // namespace rmvs

extern "C" {

RECON_MVS_API int32_t recon_mvs_hip_plane_cost_grad(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                    int32_t num_src, const int32_t* src_views,
                                                    const recon_mvs_plane_params* params,
                                                    const double* pose_delta, const float* weights,
                                                    float* cost, float* grad_plane,
                                                    double* grad_pose) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  AdjArgs g;
  int32_t st = MakeAdjArgs(ctx, ref_view, num_src, src_views, params, pose_delta, &g);
  if (st != RECON_MVS_OK) return st;
  const PlaneArgs& a = g.a;
  const size_t P = static_cast<size_t>(a.w) * a.h;
  const int tiles = ((a.w + kAdjTW - 1) / kAdjTW) * ((a.h + kAdjTH - 1) / kAdjTH);
  float* d_w = nullptr;
  if (weights) {
    d_w = static_cast<float*>(rmvs::Ensure(ctx, rmvs::kSlotAdjIn, P * sizeof(float)));
    if (!d_w) return rmvs::Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
    RMVS_HIP(ctx, hipMemcpyAsync(d_w, weights, P * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  }
  char* d_out = static_cast<char*>(rmvs::Ensure(ctx, rmvs::kSlotAdjOut, P * sizeof(float) * 5));
  double* d_part = static_cast<double*>(
      rmvs::Ensure(ctx, rmvs::kSlotAdjPart, (static_cast<size_t>(tiles) * num_src + num_src) * 6 * sizeof(double)));
  if (!d_out || !d_part) return rmvs::Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  g.weights = d_w;
  // gradients first: float4 rows stay 16-byte aligned
  g.grad_plane = grad_plane ? reinterpret_cast<float4*>(d_out) : nullptr;
  g.cost_out = reinterpret_cast<float*>(d_out + P * sizeof(float4));
  g.pose_part = grad_pose ? d_part : nullptr;
  st = LaunchAdj(ctx, params, kAdjGrad, &g);
  if (st != RECON_MVS_OK) return st;
  double* d_pose = d_part + static_cast<size_t>(tiles) * num_src * 6;
  if (grad_pose) {
    const int n = num_src * 6;
    hipLaunchKernelGGL(k_pose_reduce, dim3(1), dim3(128), 0, ctx->stream, d_part, g.a.num_tiles, n, d_pose);
    RMVS_LAUNCH_CHECK(ctx);
    RMVS_HIP(ctx, hipMemcpyAsync(grad_pose, d_pose, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  if (cost) RMVS_HIP(ctx, hipMemcpyAsync(cost, g.cost_out, P * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  if (grad_plane)
    RMVS_HIP(ctx, hipMemcpyAsync(grad_plane, g.grad_plane, P * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_plane_cost_grad_noalloc(
    recon_mvs_hip_context* ctx, int32_t ref_view, int32_t num_src, const int32_t* src_views,
    const recon_mvs_plane_params* params, const double* pose_delta, const float* d_weights,
    float* d_cost, float* d_grad_plane, double* d_grad_pose) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (reinterpret_cast<uintptr_t>(d_grad_plane) % alignof(float4) != 0)
    return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "cost_grad_noalloc: d_grad_plane not 16-byte aligned");
  if (reinterpret_cast<uintptr_t>(d_grad_pose) % alignof(double) != 0)
    return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "cost_grad_noalloc: d_grad_pose not 8-byte aligned");
  AdjArgs g;
  int32_t st = MakeAdjArgs(ctx, ref_view, num_src, src_views, params, pose_delta, &g);
  if (st != RECON_MVS_OK) return st;
  const PlaneArgs& a = g.a;
  const size_t tiles = static_cast<size_t>((a.w + kAdjTW - 1) / kAdjTW) * ((a.h + kAdjTH - 1) / kAdjTH);
  const size_t need = tiles * static_cast<size_t>(num_src) * 6 * sizeof(double);
  if (d_grad_pose && need > kNoAllocScratchBytes)
    return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT,
                      "cost_grad_noalloc: pose partials (" + std::to_string(tiles) + " tiles x " +
                          std::to_string(num_src) + " sources x 6 doubles = " + std::to_string(need) +
                          " bytes) exceed the 2 MiB scratch");
  double* part = nullptr;
  if (d_grad_pose) {
    const int dev = ctx->device;
    void* ptr = (dev >= 0 && dev < kMaxScratchDevices) ? g_scratch_ptr[dev] : nullptr;
    if (!ptr) return rmvs::Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "cost_grad_noalloc: no scratch address");
    part = static_cast<double*>(ptr);
  }
  g.weights = d_weights;
  g.grad_plane = reinterpret_cast<float4*>(d_grad_plane);
  g.cost_out = d_cost;
  g.pose_part = part;
  // every tile writes all of its partials: nothing in the scratch needs
  // clearing, and what an earlier call left there is never read
  st = LaunchAdj(ctx, params, kAdjGrad, &g, false);
  if (st != RECON_MVS_OK) return st;
  if (d_grad_pose) {
    hipLaunchKernelGGL(k_pose_reduce, dim3(1), dim3(128), 0, ctx->stream, part, g.a.num_tiles,
                       num_src * 6, d_grad_pose);
    RMVS_LAUNCH_CHECK(ctx);
  }
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_plane_polish(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                 int32_t num_src, const int32_t* src_views,
                                                 const recon_mvs_plane_params* params, int32_t steps) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (rmvs::CancelRequested(ctx)) return rmvs::Cancelled(ctx);
  if (steps < 0 || steps > 1000) return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "steps out of [0, 1000]");
  AdjArgs g;
  int32_t st = MakeAdjArgs(ctx, ref_view, num_src, src_views, params, nullptr, &g);
  if (st != RECON_MVS_OK) return st;
  const size_t P = static_cast<size_t>(g.a.w) * g.a.h;
  const double baseline = rmvs::PlaneMeanBaseline(ctx, ref_view, num_src, src_views);
  const recon_mvs_camera& R = ctx->views[ref_view]->cam;
  if (!(baseline > 1e-9)) return rmvs::Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "zero baseline");
  g.inv_per_px = static_cast<float>(1.0 / (0.5 * (R.fx + R.fy) * baseline));
  g.steps = static_cast<float2*>(rmvs::Ensure(ctx, rmvs::kSlotAdjIn, P * sizeof(float2)));
  if (!g.steps) return rmvs::Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  for (int s = 0; s < steps; ++s) {
    g.step = s;
    st = LaunchAdj(ctx, params, kAdjPolish, &g);
    if (st != RECON_MVS_OK) return st;
  }
  ctx->states[ref_view]->has_geo = false;
  return RECON_MVS_OK;
}

}  // extern "C"
