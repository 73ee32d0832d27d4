// plane_mode.hip — slanted-plane PatchMatch on MI355X (the bench path).
//
// Per pixel the state is a plane (depth D along the pixel's ray, unit normal n
// in the reference camera frame) and its aggregated cost.  A hypothesis is
// scored by warping the reference window into every source view through the
// plane-induced homography H = K_s (R_rel + t_rel m^T) K_r^-1, m = K_r^-T n / rho
// (rho = n . X0, X0 = D K_r^-1 q_c), computing a zero-mean NCC per view on a
// strided window, and averaging 1 - NCC over the best top_k views.
//
// Precision split: H and the window-centre projection are composed in FP64 per
// (hypothesis, view); the per-sample warp is applied in FP32 as an offset from
// that centre (so FP32 only carries offsets of a few pixels), bilinear taps and
// the NCC sums are FP32 (images stored as FP32, FP16 or BF16).
//
// Schedule: random plane init, then red-black half-sweeps: pixels with
// (x + y) & 1 == colour only read pixels of the other colour, so each launch
// updates half the image in place without races.  Candidates per pixel and
// iteration: 4 propagated planes (per direction, the lowest-cost other-colour
// pixel at odd distances 1..2*reach-1), 3 perturbed copies of the current
// plane, 1 random plane.  Random numbers: lowbias32 hash chain keyed by
// (seed, pass, pixel, draw) (shaders/common/gpu_algo.h:150-154).
//
// Candidate planes are generated with explicit FP32 operations and no FMA
// contraction (this file is compiled with -ffp-contract=off; the hot loops
// use fmaf() where a fused multiply-add is wanted), so the FP64 oracle can
// regenerate exactly the same candidates.
#include <type_traits>

#include "plane_common.hpp"

// Init plane of pixel (x, y): random, or the fronto-parallel plane at the
// mean depth of its seeded cell.
__device__ __forceinline__ Hyp InitPlane(const PlaneArgs& a, int x, int y) {
  float ray[3];
  PixelRay(a, x, y, ray);
  const uint32_t key = RngKey(a.seed, 0u, static_cast<uint32_t>(y * a.w + x));
  Hyp h = RandomPlane(a, key, 0, ray);
  if (a.seed_cells) {
    // a seeded cell starts on the fronto-parallel plane at its mean depth
    const float sd = a.seed_cells[min(y >> 5, a.seed_gy - 1) * a.seed_gx + min(x >> 5, a.seed_gx - 1)];
    if (sd > 0.0f) h = Hyp{sd, 0.0f, 0.0f, -1.0f};
  }
  return h;
}

// First record of pixel (x, y) in PlaneArgs::prev: 4 per pixel, the pixels
// of each colour together (colour (x + y) & 1, then the compact index
// y * ceil(w / 2) + x / 2), so a half-sweep reads and writes one dense half.
__device__ __forceinline__ int PrevSlot(const PlaneArgs& a, int x, int y) {
  const int cw = (a.w + 1) >> 1;
  return 4 * (((x + y) & 1) * a.h * cw + y * cw + (x >> 1));
}

// A new run: pixel (x, y) has been scored on no plane yet.
__device__ __forceinline__ void ClearPrev(const PlaneArgs& a, int x, int y) {
  if (!a.prev) return;
  const float n = __builtin_nanf("");
  const int s = PrevSlot(a, x, y);
#pragma unroll
  for (int k = 0; k < 4; ++k) a.prev[s + k] = make_float4(n, n, n, n);
}

// Random init + initial cost of every pixel (both colours) of one tile.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void InitTile(const PlaneArgs& a, RefTile<HALF, FMT, TW, TH>& tile,
                                         int tx, int ty) {
  const int x0 = tx * TW, y0 = ty * TH;
  tile.load(a, x0, y0);
  __syncthreads();
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    tile.pixel(x0, y0, color, &x, &y, &r);
    if (x >= a.w || y >= a.h) continue;
    RefWindow<HALF, STRIDE> win;
    tile.template window<STRIDE>(win, x, x0, r);
    const int idx = y * a.w + x;
    const Hyp h = InitPlane(a, x, y);
    int nv = 0;
    const float c = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, h, &nv, nullptr);
    a.plane[idx] = make_float4(h.d, h.nx, h.ny, h.nz);
    a.cost[idx] = c;
    a.support[idx] = static_cast<uint8_t>(min(nv, 255));
    ClearPrev(a, x, y);
  }
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_plane_init(PlaneArgs a) {
  __shared__ RefTile<HALF, FMT, TW, TH> tile;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  InitTile<HALF, STRIDE, FMT, TW, TH>(a, tile, tx, ty);
}

// Re-expresses the plane of pixel (px, py) at the ray of (x, y).
__device__ __forceinline__ bool TransferPlane(const PlaneArgs& a, float4 p, int px, int py,
                                              const float* ray, Hyp* out) {
  float rn[3];
  PixelRay(a, px, py, rn);
  const float rho = p.x * (p.y * rn[0] + p.z * rn[1] + p.w * rn[2]);
  const float dn = p.y * ray[0] + p.z * ray[1] + p.w * ray[2];
  if (!(dn < 0.0f)) return false;
  const float d = rho / dn;
  if (!(d >= a.min_depth && d <= a.max_depth)) return false;
  *out = Hyp{d, p.y, p.z, p.w};
  return true;
}

constexpr int kMaxReach = 16;

// Propagation sources of pixel (x, y): per direction the lowest-cost
// other-colour pixel at odd distances 1 .. 2*reach-1 (all reads issued
// before any comparison); -1 where there is none.
__device__ __forceinline__ void PropagationSources(const PlaneArgs& a, int x, int y, int* src_x,
                                                   int* src_y) {
  const int ddx[4] = {-1, 1, 0, 0};
  const int ddy[4] = {0, 0, -1, 1};
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    float cs[kMaxReach];
#pragma unroll
    for (int s = 0; s < kMaxReach; ++s) {
      const int nx = x + ddx[d] * (2 * s + 1), ny = y + ddy[d] * (2 * s + 1);
      const bool in = s < a.reach && nx >= 0 && ny >= 0 && nx < a.w && ny < a.h;
      cs[s] = in ? a.cost[ny * a.w + nx] : kCostInf;
    }
    src_x[d] = -1;
    src_y[d] = -1;
    float best = kCostInf;
#pragma unroll
    for (int s = 0; s < kMaxReach; ++s) {
      if (cs[s] < best) {
        best = cs[s];
        src_x[d] = x + ddx[d] * (2 * s + 1);
        src_y[d] = y + ddy[d] * (2 * s + 1);
      }
    }
  }
}

// Candidate c of a half-sweep at pixel (x, y): 0-3 the planes of the
// propagation sources, 4-6 perturbations of `base` (the plane the pixel had
// at the start of the sweep: 4 depth, 5 normal, 6 both), 7 a random plane.
// false: no candidate.
__device__ __forceinline__ bool SweepCandidate(const PlaneArgs& a, int c, const float* ray, uint32_t key,
                                               float4 base, float dscale, float nscale, const int* src_x,
                                               const int* src_y, Hyp* h) {
  bool ok = true;
  if (c < 4) {
    if (src_x[c] < 0) return false;
    ok = TransferPlane(a, a.plane[src_y[c] * a.w + src_x[c]], src_x[c], src_y[c], ray, h);
  } else if (c < 7) {
    const uint32_t k0 = 4u * static_cast<uint32_t>(c - 4);
    float d = base.x;
    float n[3] = {base.y, base.z, base.w};
    if (c != 5) {
      const float inv = 1.0f / d;
      const float u = 2.0f * RngUniform(key, k0) - 1.0f;
      const float inv2 = inv + inv * (dscale * u);
      ok = inv2 >= a.inv_min && inv2 <= a.inv_max;
      d = 1.0f / inv2;
    }
    if (ok && c != 4) {
      n[0] = n[0] + nscale * (2.0f * RngUniform(key, k0 + 1) - 1.0f);
      n[1] = n[1] + nscale * (2.0f * RngUniform(key, k0 + 2) - 1.0f);
      n[2] = n[2] + nscale * (2.0f * RngUniform(key, k0 + 3) - 1.0f);
      ok = Normalize3(n) && (n[0] * ray[0] + n[1] * ray[1] + n[2] * ray[2]) < 0.0f;
    }
    *h = Hyp{d, n[0], n[1], n[2]};
  } else {
    *h = RandomPlane(a, key, 12u, ray);
  }
  return ok;
}

// Whether hypothesis h is the plane p bit for bit.
__device__ __forceinline__ bool SamePlane(const Hyp& h, float4 p) {
  return h.d == p.x && h.nx == p.y && h.ny == p.z && h.nz == p.w;
}
__device__ __forceinline__ bool SamePlane4(float4 h, float4 p) {
  return h.x == p.x && h.y == p.y && h.z == p.z && h.w == p.w;
}

// Propagated candidates (bits of `ok`) of pixel (x, y) that need no
// scoring: equal to an earlier candidate, to the current plane `base`, or to
// a plane the pixel was scored on in an earlier half-sweep of the run
// (a.prev); then records this half-sweep's propagated candidates in a.prev.
__device__ __forceinline__ int SkipPropagated(const PlaneArgs& a, int x, int y, const float4 (&pc)[4],
                                              const bool (&ok)[4], float4 base) {
  int skip = 0;
  float4 old[4];
  const float n = __builtin_nanf("");
  const int ps = a.prev ? PrevSlot(a, x, y) : 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) old[k] = a.prev ? a.prev[ps + k] : make_float4(n, n, n, n);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    bool same = SamePlane4(pc[c], base);
#pragma unroll
    for (int k = 0; k < c; ++k) same = same || (ok[k] && SamePlane4(pc[c], pc[k]));
#pragma unroll
    for (int k = 0; k < 4; ++k) same = same || SamePlane4(pc[c], old[k]);
    if (ok[c] && same) skip |= 1 << c;
  }
  if (a.prev)
#pragma unroll
    for (int c = 0; c < 4; ++c) a.prev[ps + c] = ok[c] ? pc[c] : make_float4(n, n, n, n);
  return skip;
}

// One red-black half-sweep of one tile over the pixels of colour `color`.
//
// Candidates that are bit for bit an earlier candidate, the pixel's current
// plane, or (propagated ones) a plane the pixel was scored on in an earlier
// half-sweep of the run are skipped: they would score the same cost by the
// same operations, which lost or was taken, or equals the current cost and
// loses the strict comparison (the current cost is always the full PlaneCost
// of the current plane: the init's, or an accepted candidate's, and an
// accepted candidate never took the early exit), or, scored earlier, was
// not below the pixel's cost then, and that cost only falls during a run.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void SweepTile(const PlaneArgs& a, RefTile<HALF, FMT, TW, TH>& tile,
                                          int tx, int ty, int color, int iter) {
  const int x0 = tx * TW, y0 = ty * TH;
  tile.load(a, x0, y0);
  __syncthreads();
  int x, y, r;
  tile.pixel(x0, y0, color, &x, &y, &r);
  if (x >= a.w || y >= a.h) return;
  RefWindow<HALF, STRIDE> win;
  tile.template window<STRIDE>(win, x, x0, r);
  const int idx = y * a.w + x;
  float ray[3];
  PixelRay(a, x, y, ray);
  float4 cur = a.plane[idx];
  float ccur = a.cost[idx];
  int scur = a.support[idx];
  const uint32_t key = RngKey(a.seed, static_cast<uint32_t>(1 + 2 * iter + color),
                              static_cast<uint32_t>(idx));
  const float dscale = ldexpf(a.depth_perturb, -iter);
  const float nscale = ldexpf(a.normal_perturb, -iter);
  const float4 base = cur;
  int src_x[4], src_y[4];
  PropagationSources(a, x, y, src_x, src_y);
  // propagated candidates that need no scoring (bit c)
  int dup_mask = 0;
  {
    float4 pc[4];
    bool pok[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      Hyp h{0.0f, 0.0f, 0.0f, 0.0f};
      pok[c] = SweepCandidate(a, c, ray, key, base, dscale, nscale, src_x, src_y, &h);
      pc[c] = make_float4(h.d, h.nx, h.ny, h.nz);
    }
    dup_mask = SkipPropagated(a, x, y, pc, pok, base);
  }
  for (int c = 0; c < 8; ++c) {
    if ((dup_mask >> c) & 1) continue;
    Hyp h;
    if (!SweepCandidate(a, c, ray, key, base, dscale, nscale, src_x, src_y, &h)) continue;
    if (c >= 4 && SamePlane(h, base)) continue;
    int nv = 0;
    // candidates that provably cannot beat the current cost stop early (the
    // margin covers the rounding of the bound against the full sum)
    const float cc = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, h, &nv, nullptr,
                                                  ccur < kCostInf ? fmaf(ccur, 1.0f + 1e-5f, 1e-7f) : kCostInf);
    if (cc < ccur) {
      ccur = cc;
      cur = make_float4(h.d, h.nx, h.ny, h.nz);
      scur = nv;
    }
  }
  a.plane[idx] = cur;
  a.cost[idx] = ccur;
  a.support[idx] = static_cast<uint8_t>(min(scur, 255));
}

// Stride-2 windows: 3 waves per SIMD (at most 168 VGPRs; without the limit
// the compiler takes ~172): converged half-sweeps are VALU bound and gain from
// the third wave; the unconverged ones are held at lower occupancy by the
// per-phase caps.  Stride-1 windows would spill under the limit.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_plane_sweep(PlaneArgs a) {
  __shared__ RefTile<HALF, FMT, TW, TH> tile;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  SweepTile<HALF, STRIDE, FMT, TW, TH>(a, tile, tx, ty, a.color, a.iter);
}

// ---------------------------------------------------------------------------
// Binned evaluation: the same init and half-sweep as InitTile / SweepTile
// with the tile's hypotheses evaluated in order of inverse depth.
//
// In the unconverged phases every pixel's hypotheses lie at random depths, so
// the 64 lanes of a wave sample 64 windows scattered along as many epipolar
// segments of each source: each gather touches ~64 cache lines and nothing is
// reused across lanes.  Here the owning threads first generate their
// pixel's hypotheses (entries; those that need no scoring are dropped, as in
// SweepTile) and keep them where the scoring threads can read them (see
// BinnedLds), a counting sort in LDS puts the entries in 64 bins
// of inverse depth over [inv_min, inv_max] (the entries of pixels whose
// reference window is cut by the image border after all the others), and
// the workgroup evaluates the sorted list: a wave's 64 consecutive entries
// lie within a few bins, so their windows in every source sit in one compact
// region (the tile's extent plus a short piece of the epipolar lines) and
// share cache lines, and the window loops that check reference samples stay
// in the waves of border pixels.  Evaluating the list also compacts the
// work: duplicate and invalid candidates leave no idle lanes.  The owning
// threads then pick each pixel's best entry in candidate order.
//
// Results are bit-identical to InitTile / SweepTile: every entry is scored
// by the same PlaneCost operations (a pixel's result does not depend on
// which lanes share its wave, see ViewNcc); the early exit rejects against
// the pixel's best cost so far (its cost at the start of the sweep, lowered
// by each of its entries scored in full, in whatever order the waves finish
// them), so an entry dropped there costs more than the start cost or than
// another of the pixel's entries and is not the pick, as with SweepTile's
// running best; and the choice runs in candidate order with the same strict
// comparison.
// ---------------------------------------------------------------------------
constexpr int kBins = 64;              // inverse-depth bins per pixel class
constexpr int kBinnedEntries = 2048;   // sweep: 256 pixels x 8 candidates; init: 512 pixels x 1

// LDS of the binned forms (about 45 KB: 3 workgroups per CU).  Entries hold
// no hypothesis: the propagated ones (candidates 0-3) are kept here per
// pixel, the perturbed and random ones (4-7) in the pixel's slots of
// PlaneArgs::cand and the init planes in the plane array, each written by
// the owning thread before the sort and read by the scoring thread after it
// (the workgroup barriers order the global writes and reads).
template <int HALF, int FMT, int TW, int TH>
struct BinnedLds {
  RefTile<HALF, FMT, TW, TH> tile;
  float4 prop[256 * 4];            // sweep: pixel t's propagated candidate c at [4 t + c]
  float res[kBinnedEntries];       // cost of entry e (kCostInf: rejected)
  uint8_t nv[kBinnedEntries];      // its valid views
  uint16_t list[kBinnedEntries];   // entries in bin order
  float4 wstat[512];               // per pixel slot: the reference window's sum_c, sum_c2, n, mean
  int best_bits[256];              // sweep: float bits of each pixel's best cost so far (its cost at
                                   // the start of the sweep, lowered as its entries are scored)
  int bin[2 * kBins];              // counts, then first positions (whole windows first, then by depth)
  int n;                           // entries in the list
};

// Slot of candidate c (4-7) of pixel (x, y) in PlaneArgs::cand: 4 slots per
// pixel of the active colour, by its compact index y * ceil(w / 2) + x / 2
// (the colour's pixels of a row have distinct x / 2).
__device__ __forceinline__ int CandSlot(const PlaneArgs& a, int x, int y, int c) {
  return 4 * (y * ((a.w + 1) >> 1) + (x >> 1)) + (c - 4);
}

// Sort bin of a hypothesis at depth d of a pixel whose reference window is
// whole (`full`) or not: the bin of 1/d in [inv_min, inv_max] (clamped; NaN
// -> 0), after all kBins of the whole windows when not.
__device__ __forceinline__ int SortBin(const PlaneArgs& a, float d, bool full) {
  const float u = (1.0f / d - a.inv_min) * (static_cast<float>(kBins) / (a.inv_max - a.inv_min));
  return static_cast<int>(fminf(fmaxf(u, 0.0f), static_cast<float>(kBins - 1))) + (full ? 0 : kBins);
}

// Counting sort of the workgroup's entries by bin.  Each thread passes its
// NE entries as (bin << 16 | entry), -1 for none; on return L.list holds
// them bin by bin and L.n their count.  Inside a bin a thread's entries are
// consecutive (one reservation per distinct bin of the thread), so a wave's
// lanes cover a few neighbouring pixels with all their candidates at that
// depth rather than one candidate of 64 pixels (order across threads is the
// order of the atomics, otherwise unspecified).
template <int NE, int HALF, int FMT, int TW, int TH>
__device__ __forceinline__ void BinEntries(BinnedLds<HALF, FMT, TW, TH>& L, const int (&be)[NE]) {
  const int t = threadIdx.x;
  if (t < 2 * kBins) L.bin[t] = 0;
  __syncthreads();
  // one atomic per distinct bin of the thread, for all its entries there:
  // a pixel's entries in a bin stay consecutive
  int rank[NE];
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int bk = be[k] >> 16;
    int before = 0, cnt = 0;
#pragma unroll
    for (int m = 0; m < NE; ++m) {
      const bool same = be[m] >= 0 && (be[m] >> 16) == bk;
      before += (m < k && same) ? 1 : 0;
      cnt += same ? 1 : 0;
    }
    rank[k] = (be[k] >= 0 && before == 0) ? atomicAdd(&L.bin[bk], cnt) : 0;
  }
#pragma unroll
  for (int k = 0; k < NE; ++k) {
    const int bk = be[k] >> 16;
    int before = 0, base = 0;
#pragma unroll
    for (int m = 0; m < NE; ++m) {
      const bool same = be[m] >= 0 && (be[m] >> 16) == bk;
      if (m < k && same) {
        base = before == 0 ? rank[m] : base;
        ++before;
      }
    }
    if (be[k] >= 0 && before > 0) rank[k] = base + before;
  }
  __syncthreads();
  if (t == 0) {
    int s = 0;
    for (int b = 0; b < 2 * kBins; ++b) {
      const int c = L.bin[b];
      L.bin[b] = s;
      s += c;
    }
    L.n = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NE; ++k)
    if (be[k] >= 0) L.list[L.bin[be[k] >> 16] + rank[k]] = static_cast<uint16_t>(be[k] & 0xFFFF);
  __syncthreads();
}

// Reference window of pixel slot s from its stored sums.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void SlotWindow(const BinnedLds<HALF, FMT, TW, TH>& L, int s, int x, int x0, int r,
                                           RefWindow<HALF, STRIDE>& win) {
  using Tile = RefTile<HALF, FMT, TW, TH>;
  const float4 ws = L.wstat[s];
  win.base = &L.tile.v[r][x - x0];
  win.rs = STRIDE * Tile::LW;
  win.shift = L.tile.shift;
  win.sum_c = ws.x;
  win.sum_c2 = ws.y;
  win.n = ws.z;
  win.mean = ws.w;
  win.full = (ws.z == static_cast<float>(RefWindow<HALF, STRIDE>::S));
}

// Scores the sorted entries.  Init: entry e = pixel slot e (colour e >> 8,
// thread e & 255); sweep: entry e = candidate e & 7 of thread e >> 3's pixel,
// with early exit against the pixel's best cost so far.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void ScoreEntries(const PlaneArgs& a, BinnedLds<HALF, FMT, TW, TH>& L, int x0,
                                             int y0, int color, bool sweep) {
  const int n = L.n;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int e = L.list[i];
    const int s = sweep ? (e >> 3) : e;
    const int col = sweep ? color : (s >> 8);
    int x, y, r;
    L.tile.pixel_of(s & 255, x0, y0, col, &x, &y, &r);
    RefWindow<HALF, STRIDE> win;
    SlotWindow<HALF, STRIDE, FMT, TW, TH>(L, s, x, x0, r, win);
    float reject_at = kCostInf;
    float4 p;
    if (sweep) {
      reject_at = RejectAbove(__int_as_float(L.best_bits[s]));
      const int c = e & 7;
      p = c < 4 ? L.prop[4 * s + c] : a.cand[CandSlot(a, x, y, c)];
    } else {
      p = a.plane[y * a.w + x];  // the init plane, written by the owning thread
    }
    const Hyp h{p.x, p.y, p.z, p.w};
    int nv = 0;
    const float cost = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, h, &nv, nullptr, reject_at,
                                                    sweep ? &L.best_bits[s] : nullptr);
    // a full cost bounds the pixel's pick from above: the pixel's entries
    // scored after it exit against it (exact: one that exits costs more)
    if (sweep && cost < kCostInf) atomicMin(&L.best_bits[s], __float_as_int(cost));
    L.res[e] = cost;
    L.nv[e] = static_cast<uint8_t>(min(nv, 255));
  }
  __syncthreads();
}

// InitTile with binned evaluation.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void InitTileBinned(const PlaneArgs& a, BinnedLds<HALF, FMT, TW, TH>& L, int tx,
                                               int ty) {
  const int x0 = tx * TW, y0 = ty * TH;
  L.tile.load(a, x0, y0);
  __syncthreads();
  int be[2] = {-1, -1};
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    L.tile.pixel(x0, y0, color, &x, &y, &r);
    if (x >= a.w || y >= a.h) continue;
    const int s = color * 256 + static_cast<int>(threadIdx.x);
    RefWindow<HALF, STRIDE> win;
    L.tile.template window<STRIDE>(win, x, x0, r);
    L.wstat[s] = make_float4(win.sum_c, win.sum_c2, win.n, win.mean);
    const Hyp h = InitPlane(a, x, y);
    a.plane[y * a.w + x] = make_float4(h.d, h.nx, h.ny, h.nz);
    be[color] = (SortBin(a, h.d, win.full) << 16) | s;
  }
  BinEntries(L, be);
  ScoreEntries<HALF, STRIDE, FMT, TW, TH>(a, L, x0, y0, 0, false);
  for (int color = 0; color < 2; ++color) {
    if (be[color] < 0) continue;
    int x, y, r;
    L.tile.pixel(x0, y0, color, &x, &y, &r);
    const int s = be[color] & 0xFFFF;
    const int idx = y * a.w + x;
    a.cost[idx] = L.res[s];
    a.support[idx] = L.nv[s];
    ClearPrev(a, x, y);
  }
}

// SweepTile with binned evaluation.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__device__ __forceinline__ void SweepTileBinned(const PlaneArgs& a, BinnedLds<HALF, FMT, TW, TH>& L, int tx,
                                                int ty, int color, int iter) {
  const int x0 = tx * TW, y0 = ty * TH;
  L.tile.load(a, x0, y0);
  __syncthreads();
  const int t = threadIdx.x;
  int x, y, r;
  L.tile.pixel(x0, y0, color, &x, &y, &r);
  const bool in = x < a.w && y < a.h;
  int be[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) be[c] = -1;
  int idx = 0;
  float4 cur = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float ccur = kCostInf;
  int scur = 0;
  if (in) {
    RefWindow<HALF, STRIDE> win;
    L.tile.template window<STRIDE>(win, x, x0, r);
    L.wstat[t] = make_float4(win.sum_c, win.sum_c2, win.n, win.mean);
    idx = y * a.w + x;
    float ray[3];
    PixelRay(a, x, y, ray);
    cur = a.plane[idx];
    ccur = a.cost[idx];
    scur = a.support[idx];
    L.best_bits[t] = __float_as_int(ccur);
    const uint32_t key = RngKey(a.seed, static_cast<uint32_t>(1 + 2 * iter + color),
                                static_cast<uint32_t>(idx));
    const float dscale = ldexpf(a.depth_perturb, -iter);
    const float nscale = ldexpf(a.normal_perturb, -iter);
    const float4 base = cur;
    int src_x[4], src_y[4];
    PropagationSources(a, x, y, src_x, src_y);
    float4 pc[4];
    bool pok[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      Hyp h{0.0f, 0.0f, 0.0f, 0.0f};
      pok[c] = SweepCandidate(a, c, ray, key, base, dscale, nscale, src_x, src_y, &h);
      pc[c] = make_float4(h.d, h.nx, h.ny, h.nz);
      L.prop[4 * t + c] = pc[c];
    }
    // candidates that need no scoring are skipped, as in SweepTile
    const int dup_mask = SkipPropagated(a, x, y, pc, pok, base);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      if ((dup_mask >> c) & 1) continue;
      Hyp h;
      if (c < 4) {
        if (!pok[c]) continue;
        h = Hyp{pc[c].x, pc[c].y, pc[c].z, pc[c].w};
      } else if (!SweepCandidate(a, c, ray, key, base, dscale, nscale, src_x, src_y, &h) ||
                 SamePlane(h, base)) {
        continue;
      }
      // PlaneCost scores these kCostInf without any view: never taken
      if (!(h.d >= a.min_depth && h.d <= a.max_depth) || !MakeHypGeom(a, x, y, h).front) continue;
      if (c >= 4) a.cand[CandSlot(a, x, y, c)] = make_float4(h.d, h.nx, h.ny, h.nz);
      be[c] = (SortBin(a, h.d, win.full) << 16) | (t * 8 + c);
    }
  }
  BinEntries(L, be);
  ScoreEntries<HALF, STRIDE, FMT, TW, TH>(a, L, x0, y0, color, true);
  if (!in) return;
  int best = -1;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    if (be[c] < 0) continue;
    const float cc = L.res[t * 8 + c];
    if (cc < ccur) {
      ccur = cc;
      scur = L.nv[t * 8 + c];
      best = c;
    }
  }
  if (best >= 4) cur = a.cand[CandSlot(a, x, y, best)];
  else if (best >= 0) cur = L.prop[4 * t + best];
  a.plane[idx] = cur;
  a.cost[idx] = ccur;
  a.support[idx] = static_cast<uint8_t>(min(scur, 255));
}

// Single-view launches of the binned forms (3 workgroups per CU, as the
// fused forms).
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_plane_init_binned(PlaneArgs a) {
  __shared__ BinnedLds<HALF, FMT, TW, TH> L;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  InitTileBinned<HALF, STRIDE, FMT, TW, TH>(a, L, tx, ty);
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_plane_sweep_binned(PlaneArgs a) {
  __shared__ BinnedLds<HALF, FMT, TW, TH> L;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  SweepTileBinned<HALF, STRIDE, FMT, TW, TH>(a, L, tx, ty, a.color, a.iter);
}

// ---------------------------------------------------------------------------
// View-phased half-sweep: the same half-sweep as k_plane_sweep in 2 + nsrc
// launches.  k_phase_prep makes each pixel's 8 candidates, one k_phase_view
// launch per source view adds that view to every live candidate's
// accumulators (in memory between launches), and k_phase_final scores the
// candidates and keeps the best in candidate order.  Each view launch touches
// one source image only, and visits the tiles in strips along that view's
// epipolar lines (`order`), so the source footprint of the tiles in flight on
// an XCD stays within its L2 -- the unconverged sweeps' gathers otherwise
// spread over all sources at once.
//
// Results are bit-identical to k_plane_sweep: every candidate's view costs
// are accumulated by the same operations in the same view order; the early
// exit uses the cost the pixel had at the start of the sweep, a bound no
// tighter than the running best of k_plane_sweep, so it drops only candidates
// that k_plane_sweep drops too or that lose to the running best anyway.
// ---------------------------------------------------------------------------

// Candidate state in the colour's compact pixel index ci = y * cw + (x >> 1),
// structure of arrays, slot c of candidate c: [c * n + ci].
struct PhaseState {
  float4* cand;  // candidate planes
  float* topk;   // [(c * K + k) * n + ci]: the K smallest view costs so far
  float* csum;   // sum of the view costs so far
  int* nvalid;   // valid views so far; -1: candidate out (invalid or rejected)
  int n, cw, K;  // K = top_k when the k smallest are aggregated, else 0
};

// The pixel of a thread in a tile, as RefTile::pixel.
template <int TW>
__device__ __forceinline__ void TilePixel(int x0, int y0, int color, int* x, int* y) {
  const int r = threadIdx.x / (TW / 2);
  const int j = threadIdx.x % (TW / 2);
  *y = y0 + r;
  *x = x0 + 2 * j + (((x0 + *y) & 1) ^ color);
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) void k_phase_prep(PlaneArgs a, PhaseState st, int color, int iter) {
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  int x, y;
  TilePixel<TW>(tx * TW, ty * TH, color, &x, &y);
  if (x >= a.w || y >= a.h) return;
  const int idx = y * a.w + x;
  const int ci = y * st.cw + (x >> 1);
  float ray[3];
  PixelRay(a, x, y, ray);
  const float4 base = a.plane[idx];
  const uint32_t key = RngKey(a.seed, static_cast<uint32_t>(1 + 2 * iter + color),
                              static_cast<uint32_t>(idx));
  const float dscale = ldexpf(a.depth_perturb, -iter);
  const float nscale = ldexpf(a.normal_perturb, -iter);
  int src_x[4], src_y[4];
  PropagationSources(a, x, y, src_x, src_y);
  for (int c = 0; c < 8; ++c) {
    Hyp h{0.0f, 0.0f, 0.0f, -1.0f};
    // a candidate PlaneCost would score kCostInf without any view is out now
    const bool ok = SweepCandidate(a, c, ray, key, base, dscale, nscale, src_x, src_y, &h) &&
                    h.d >= a.min_depth && h.d <= a.max_depth && MakeHypGeom(a, x, y, h).front;
    const size_t i = static_cast<size_t>(c) * st.n + ci;
    st.cand[i] = make_float4(h.d, h.nx, h.ny, h.nz);
    st.nvalid[i] = ok ? 0 : -1;
    st.csum[i] = 0.0f;
    for (int k = 0; k < st.K; ++k) st.topk[(static_cast<size_t>(c) * st.K + k) * st.n + ci] = kCostInf;
  }
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_phase_view(PlaneArgs a, PhaseState st, const int* __restrict__ order, int v, int color) {
  __shared__ RefTile<HALF, FMT, TW, TH> tile;
  const int b = blockIdx.x;
  const int t = (b & 7) * a.tiles_per_xcd + (b >> 3);  // XCD-aware, as TileOf
  if (t >= a.num_tiles) return;
  const int packed = order[t];
  const int x0 = (packed & 0xFFFF) * TW, y0 = (packed >> 16) * TH;
  tile.load(a, x0, y0);
  __syncthreads();
  int x, y, r;
  tile.pixel(x0, y0, color, &x, &y, &r);
  if (x >= a.w || y >= a.h) return;
  RefWindow<HALF, STRIDE> win;
  tile.template window<STRIDE>(win, x, x0, r);
  const int ci = y * st.cw + (x >> 1);
  const float ccur = a.cost[y * a.w + x];
  const float reject_at = ccur < kCostInf ? fmaf(ccur, 1.0f + 1e-5f, 1e-7f) : kCostInf;
  const bool use_topk = a.top_k < a.nsrc;
  for (int c = 0; c < 8; ++c) {
    const size_t i = static_cast<size_t>(c) * st.n + ci;
    const int nv = st.nvalid[i];
    if (nv < 0) continue;
    const float4 p = st.cand[i];
    const HypGeom hg = MakeHypGeom(a, x, y, Hyp{p.x, p.y, p.z, p.w});
    CostAcc acc;
    acc.nvalid = nv;
    acc.csum = st.csum[i];
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k)
      acc.topk[k] = k < st.K ? st.topk[(static_cast<size_t>(c) * st.K + k) * st.n + ci] : kCostInf;
    (void)ViewNcc<HALF, STRIDE, FMT>(a, v, hg, win, use_topk, acc);
    if (reject_at < kCostInf && CostLowerBound(a, v, acc) > reject_at) {
      st.nvalid[i] = -1;
      continue;
    }
    if (acc.nvalid == nv) continue;  // view v invalid for it: nothing changed
    st.nvalid[i] = acc.nvalid;
    st.csum[i] = acc.csum;
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k)
      if (k < st.K) st.topk[(static_cast<size_t>(c) * st.K + k) * st.n + ci] = acc.topk[k];
  }
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) void k_phase_final(PlaneArgs a, PhaseState st, int color) {
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  int x, y;
  TilePixel<TW>(tx * TW, ty * TH, color, &x, &y);
  if (x >= a.w || y >= a.h) return;
  const int idx = y * a.w + x;
  const int ci = y * st.cw + (x >> 1);
  float4 cur = a.plane[idx];
  float ccur = a.cost[idx];
  int scur = a.support[idx];
  const bool use_topk = a.top_k < a.nsrc;
  for (int c = 0; c < 8; ++c) {
    const size_t i = static_cast<size_t>(c) * st.n + ci;
    CostAcc acc;
    acc.nvalid = st.nvalid[i];
    if (acc.nvalid < 0) continue;
    acc.csum = st.csum[i];
#pragma unroll
    for (int k = 0; k < kMaxTopKNet; ++k)
      acc.topk[k] = k < st.K ? st.topk[(static_cast<size_t>(c) * st.K + k) * st.n + ci] : kCostInf;
    const float cc = FinishCost(a, acc, use_topk);
    if (cc < ccur) {
      ccur = cc;
      cur = st.cand[i];
      scur = acc.nvalid;
    }
  }
  a.plane[idx] = cur;
  a.cost[idx] = ccur;
  a.support[idx] = static_cast<uint8_t>(min(scur, 255));
}

// Batched, persistent form: one launch processes (job, tile) units of many
// reference views of any size.  phase < 0: random init; otherwise the
// half-sweep of colour phase & 1 of iteration phase >> 1.
//
// The units are split into one queue per XCD (queue q: units [qoff[q],
// qoff[q + 1])), each holding a vertical band of every job's tiles, in
// column order (the layout TileOf gives a single view).  Workgroup ids are
// dealt round-robin to the XCDs, so workgroup b serves queue b & 7 -- the
// tiles in flight on an XCD form a compact block whose source footprints
// stay in that XCD's L2 -- and, once that queue is empty, the others in turn.
// Every workgroup exits after finding each queue empty once (the counters
// only grow).  `cancel` (the context's cancel flag in coherent host memory;
// may be null) is read by thread 0 each time it takes a unit and obeyed at
// its next take (the read overlaps the unit's work): once the flag is set,
// every workgroup stops after at most one more unit.
constexpr int kXcds = 8;
constexpr int kCounterStride = 32;  // queue counters 128 bytes apart

template <bool BINNED, int HALF, int FMT, int TW, int TH>
using BatchLds = std::conditional_t<BINNED, BinnedLds<HALF, FMT, TW, TH>, RefTile<HALF, FMT, TW, TH>>;

template <int HALF, int STRIDE, int FMT, int TW, int TH, bool BINNED>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(STRIDE == 2 ? 3 : 1)))
void k_plane_batch(const PlaneArgs* __restrict__ args, const int2* __restrict__ units,
                   const int* __restrict__ qoff, unsigned int* __restrict__ counters, int phase,
                   const int32_t* cancel) {
  __shared__ BatchLds<BINNED, HALF, FMT, TW, TH> lds;
  __shared__ int s_unit;
  int stop = 0;    // thread 0: the flag as read at the previous take
  int tried = 0;   // thread 0: queues found empty (own queue first)
  for (;;) {
    if (threadIdx.x == 0) {
      int u = -1;
      while (!stop && tried < kXcds) {
        const int q = (static_cast<int>(blockIdx.x) + tried) & (kXcds - 1);
        const int k = static_cast<int>(atomicAdd(&counters[q * kCounterStride], 1u));
        if (qoff[q] + k < qoff[q + 1]) {
          u = qoff[q] + k;
          break;
        }
        ++tried;
      }
      s_unit = u;
      if (cancel) stop = __hip_atomic_load(cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
    __syncthreads();
    const int u = __builtin_amdgcn_readfirstlane(s_unit);
    if (u < 0) return;
    const int2 un = units[u];
    const int job = __builtin_amdgcn_readfirstlane(un.x);
    const int t = __builtin_amdgcn_readfirstlane(un.y);
    const PlaneArgs& a = args[job];
    const int tx = t % a.tiles_x, ty = t / a.tiles_x;
    if constexpr (BINNED) {
      if (phase < 0) InitTileBinned<HALF, STRIDE, FMT, TW, TH>(a, lds, tx, ty);
      else SweepTileBinned<HALF, STRIDE, FMT, TW, TH>(a, lds, tx, ty, phase & 1, phase >> 1);
    } else {
      if (phase < 0) InitTile<HALF, STRIDE, FMT, TW, TH>(a, lds, tx, ty);
      else SweepTile<HALF, STRIDE, FMT, TW, TH>(a, lds, tx, ty, phase & 1, phase >> 1);
    }
    __syncthreads();  // the LDS and s_unit are reused by the next unit
  }
}

// Clears the batch's queue counters (a kernel, not a memset: the launches
// are captured into a graph).
__global__ void k_reset_counters(unsigned int* c) {
  if (threadIdx.x < kXcds) c[threadIdx.x * kCounterStride] = 0u;
}

// Scores a given plane state (both colours).
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) void k_plane_score(PlaneArgs a) {
  __shared__ RefTile<HALF, FMT, TW, TH> tile;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  const int x0 = tx * TW, y0 = ty * TH;
  tile.load(a, x0, y0);
  __syncthreads();
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    tile.pixel(x0, y0, color, &x, &y, &r);
    if (x >= a.w || y >= a.h) continue;
    RefWindow<HALF, STRIDE> win;
    tile.template window<STRIDE>(win, x, x0, r);
    const int idx = y * a.w + x;
    const float4 p = a.plane[idx];
    int nv = 0;
    const float c = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, Hyp{p.x, p.y, p.z, p.w}, &nv, nullptr);
    a.cost[idx] = c;
    a.support[idx] = static_cast<uint8_t>(min(nv, 255));
  }
}

// Sub-pixel refinement along each pixel's ray with its normal fixed.
template <int HALF, int STRIDE, int FMT, int TW, int TH>
__global__ __launch_bounds__(256) void k_plane_refine(PlaneArgs a, float inv_per_px, int iterations) {
  __shared__ RefTile<HALF, FMT, TW, TH> tile;
  int tx, ty;
  if (!TileOf(a, &tx, &ty)) return;
  const int x0 = tx * TW, y0 = ty * TH;
  tile.load(a, x0, y0);
  __syncthreads();
  for (int color = 0; color < 2; ++color) {
    int x, y, r;
    tile.pixel(x0, y0, color, &x, &y, &r);
    if (x >= a.w || y >= a.h) continue;
    const int idx = y * a.w + x;
    const float4 p = a.plane[idx];
    float best_c = a.cost[idx];
    if (!(best_c < kCostInf)) continue;
    RefWindow<HALF, STRIDE> win;
    tile.template window<STRIDE>(win, x, x0, r);
    int best_s = a.support[idx];
    const float inv0 = 1.0f / p.x;
    float best_i = inv0;
    auto eval = [&](float inv, int* nv) -> float {
      if (!(inv >= a.inv_min && inv <= a.inv_max)) {
        *nv = 0;
        return kCostInf;
      }
      return PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, Hyp{1.0f / inv, p.y, p.z, p.w}, nv, nullptr);
    };
    const float step = 0.5f * inv_per_px;
    float f[5];
    int nvs[5];
    bool fit_ok = true;
    for (int k = 0; k < 5; ++k) {
      if (k == 2) {
        f[k] = best_c;
        nvs[k] = best_s;
        continue;
      }
      f[k] = eval(inv0 + static_cast<float>(k - 2) * step, &nvs[k]);
      fit_ok = fit_ok && f[k] < kCostInf;
    }
    if (fit_ok) {
      // least-squares parabola through 5 equally spaced samples (x = -2..2)
      const float aa = (2.0f * f[0] - f[1] - 2.0f * f[2] - f[3] + 2.0f * f[4]) / 14.0f;
      const float bb = (-2.0f * f[0] - f[1] + f[3] + 2.0f * f[4]) / 10.0f;
      if (aa > 1e-12f) {
        const float vtx = -bb / (2.0f * aa);
        if (vtx >= -2.0f && vtx <= 2.0f) {
          int nv = 0;
          const float cand = inv0 + vtx * step;
          const float cc = eval(cand, &nv);
          if (cc < best_c) {
            best_c = cc;
            best_i = cand;
            best_s = nv;
          }
        }
      }
    }
    float lo = fmaxf(best_i - inv_per_px, a.inv_min), hi = fminf(best_i + inv_per_px, a.inv_max);
    if (hi > lo && iterations > 0) {
      const float ip = 0.6180339887498949f;
      float c1 = hi - ip * (hi - lo), c2 = lo + ip * (hi - lo);
      int n1 = 0, n2 = 0;
      float f1 = eval(c1, &n1), f2 = eval(c2, &n2);
      for (int k = 0; k < iterations; ++k) {
        // each lane moves its bracket, then one evaluation serves both sides
        // (lanes that went different ways would otherwise run it once per
        // side); the same points in the same order
        const bool left = f1 < f2;
        float cn;
        if (left) {
          hi = c2; c2 = c1; f2 = f1; n2 = n1;
          cn = hi - ip * (hi - lo);
        } else {
          lo = c1; c1 = c2; f1 = f2; n1 = n2;
          cn = lo + ip * (hi - lo);
        }
        int nn = 0;
        const float fn = eval(cn, &nn);
        if (left) {
          c1 = cn; f1 = fn; n1 = nn;
        } else {
          c2 = cn; f2 = fn; n2 = nn;
        }
      }
      if (f1 < best_c) { best_c = f1; best_i = c1; best_s = n1; }
      if (f2 < best_c) { best_c = f2; best_i = c2; best_s = n2; }
    }
    a.plane[idx] = make_float4(1.0f / best_i, p.y, p.z, p.w);
    a.cost[idx] = best_c;
    a.support[idx] = static_cast<uint8_t>(min(best_s, 255));
  }
}

// Cost of explicit hypotheses (parity tests): one thread per query, 64 per
// workgroup; each thread copies its reference window into LDS first.
template <int HALF, int STRIDE, int FMT>
__global__ __launch_bounds__(64) void k_plane_eval(PlaneArgs a, int n, const float* __restrict__ hyps,
                                                   float* __restrict__ cost, int* __restrict__ nvalid,
                                                   float* __restrict__ view_ncc) {
  constexpr int WS = 2 * HALF + 1;
  __shared__ float wbuf[64][WS * WS];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const float* hq = hyps + 6 * q;
  const int x = static_cast<int>(hq[0]), y = static_cast<int>(hq[1]);
  if (x < 0 || y < 0 || x >= a.w || y >= a.h) {
    cost[q] = kCostInf;
    nvalid[q] = 0;
    if (view_ncc)
      for (int v = 0; v < a.nsrc; ++v) view_ncc[q * a.nsrc + v] = __builtin_nanf("");
    return;
  }
  float* mine = wbuf[threadIdx.x];
  const float sh = Pix<FMT>(a.ref_img, y * a.ref_pitch + x);
  for (int dy = -HALF; dy <= HALF; ++dy)
    for (int dx = -HALF; dx <= HALF; ++dx) {
      const int gx = x + dx, gy = y + dy;
      mine[(dy + HALF) * WS + dx + HALF] = (gx < 0 || gy < 0 || gx >= a.w || gy >= a.h)
                                               ? __builtin_nanf("")
                                               : Pix<FMT>(a.ref_img, gy * a.ref_pitch + gx) - sh;
    }
  RefWindow<HALF, STRIDE> win;
  MakeRefWindow<HALF, STRIDE>(win, mine, STRIDE * WS, sh);
  float vn[kMaxViews];
  int nv = 0;
  const float c = PlaneCost<HALF, STRIDE, FMT>(a, x, y, win, Hyp{hq[2], hq[3], hq[4], hq[5]}, &nv,
                                               view_ncc ? vn : nullptr);
  cost[q] = c;
  nvalid[q] = nv;
  if (view_ncc)
    for (int v = 0; v < a.nsrc; ++v) view_ncc[q * a.nsrc + v] = vn[v];
}

__device__ __forceinline__ uint16_t ToBf16(float v) {
  // round to nearest even into bfloat16
  const uint32_t b = __float_as_uint(v);
  return static_cast<uint16_t>((b + 0x7FFFu + ((b >> 16) & 1u)) >> 16);
}

// Converts an uploaded view (u8 or f32 rows) into the row-pair storage layout
// (see BufQuad): one thread per element.
// BT.601 gray of one B,G,R,A pixel in integers, as BgraToGray
// (core/src/orchestrator/orchestrator.cpp:53-67); the weights sum to 256.
__device__ __forceinline__ int GrayOf(const uint8_t* px) {
  return (77 * static_cast<int>(px[2]) + 150 * static_cast<int>(px[1]) + 29 * static_cast<int>(px[0])) >> 8;
}

__global__ void k_bgra_to_gray(const uint8_t* __restrict__ bgra, int w, int h, int stride,
                               uint8_t* __restrict__ gray) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w * h) return;
  const int y = i / w, x = i % w;
  gray[i] = static_cast<uint8_t>(GrayOf(bgra + static_cast<size_t>(y) * stride + 4 * x));
}

template <int FMT>
__global__ void k_store_view(const void* __restrict__ src, int src_format, int src_stride, int w,
                             int h, int pitch, void* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pitch * h) return;
  const int y = i / pitch, x = i % pitch;
  auto at = [&](int xx, int yy) -> float {
    if (xx >= w || yy >= h) return 0.0f;
    if (src_format == RECON_MVS_PIXEL_BGRA8)
      return static_cast<float>(
          GrayOf(static_cast<const uint8_t*>(src) + static_cast<size_t>(yy) * src_stride + 4 * xx));
    return src_format == RECON_MVS_PIXEL_U8
               ? static_cast<float>(static_cast<const uint8_t*>(src)[yy * w + xx])
               : static_cast<const float*>(src)[yy * w + xx];
  };
  const float v0 = at(x, y), v1 = at(x, y + 1);
  if constexpr (FMT == RECON_MVS_PIXEL_F32) {
    static_cast<float2*>(dst)[i] = make_float2(v0, v1);
  } else if constexpr (FMT == RECON_MVS_PIXEL_F16) {
    _Float16* o = static_cast<_Float16*>(dst) + 2 * static_cast<size_t>(i);
    o[0] = static_cast<_Float16>(v0);
    o[1] = static_cast<_Float16>(v1);
  } else if constexpr (FMT == RECON_MVS_PIXEL_U8) {
    // integer values 0..255 (set_views admits only 8-bit views)
    uint8_t* o = static_cast<uint8_t*>(dst) + 2 * static_cast<size_t>(i);
    o[0] = static_cast<uint8_t>(v0);
    o[1] = static_cast<uint8_t>(v1);
  } else {
    uint16_t* o = static_cast<uint16_t*>(dst) + 2 * static_cast<size_t>(i);
    o[0] = ToBf16(v0);
    o[1] = ToBf16(v1);
  }
}

// Geometric consistency of map r against the other listed maps (FP64).
struct GeoView {
  int w, h;
  const float4* plane;
  const float* cost;
  recon_mvs_camera cam;
};
constexpr int kMaxGeo = 16;
struct GeoArgs {
  int nmaps, self;
  GeoView v[kMaxGeo];
  double thr;
  uint8_t* geo_support;
};

__global__ void k_geo_consistency(GeoArgs g) {
  const GeoView& me = g.v[g.self];
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= me.w * me.h) return;
  const float c = me.cost[idx];
  const float d = me.plane[idx].x;
  int support = 0;
  if (c < kCostInf) {
    const int x = idx % me.w, y = idx / me.w;
    double ray[3];
    PixelToRay(me.cam, x + 0.5, y + 0.5, ray);
    const double dd = static_cast<double>(d);
    const double pc[3] = {ray[0] * dd, ray[1] * dd, ray[2] * dd};
    double pw[3];
    CamToWorld(me.cam, pc, pw);
    support = 1;
    for (int m = 0; m < g.nmaps; ++m) {
      if (m == g.self) continue;
      const GeoView& o = g.v[m];
      double ps[3];
      WorldToCam(o.cam, pw, ps);
      if (ps[2] <= 1e-6) continue;
      double u, v;
      if (!Project(o.cam, ps, &u, &v)) continue;
      const double fu = floor(u), fv = floor(v);
      if (fu < 0.0 || fv < 0.0 || fu >= o.w || fv >= o.h) continue;
      const int q = static_cast<int>(fv) * o.w + static_cast<int>(fu);
      if (!(o.cost[q] < kCostInf)) continue;
      const double zm = static_cast<double>(o.plane[q].x);
      const double e = fabs(ps[2] - zm) / fmax(ps[2], zm);
      if (e <= g.thr) ++support;
    }
  }
  g.geo_support[idx] = static_cast<uint8_t>(min(support, 255));
}

// Output of one slot (finalize): depth NaN where invalid or filtered.
__global__ void k_plane_output(int P, const float4* __restrict__ plane, const float* __restrict__ cost,
                               const uint8_t* __restrict__ geo, int min_views, float* __restrict__ depth,
                               float* __restrict__ conf) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float c = cost[i];
  bool ok = c < kCostInf;
  if (geo) ok = ok && geo[i] >= min_views;
  depth[i] = ok ? plane[i].x : __builtin_nanf("");
  conf[i] = ok ? 1.0f - fminf(fmaxf(c, 0.0f), 1.0f) : 0.0f;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

namespace {

void FreeView(ResidentView* v) {
  if (v && v->pixels) (void)hipFree(v->pixels);
  delete v;
}

void FreeState(PlaneState* s) {
  if (!s) return;
  if (s->plane) (void)hipFree(s->plane);
  if (s->prev) (void)hipFree(s->prev);
  if (s->cand) (void)hipFree(s->cand);
  if (s->cost) (void)hipFree(s->cost);
  if (s->support) (void)hipFree(s->support);
  if (s->geo_support) (void)hipFree(s->geo_support);
  if (s->seed_cells) (void)hipFree(s->seed_cells);
  delete s;
}

int32_t CheckParams(recon_mvs_hip_context* ctx, const recon_mvs_plane_params* p) {
  if (!p) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null plane params");
  if (!(p->min_depth_m > 0.0 && p->max_depth_m > p->min_depth_m))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad depth range");
  if (p->top_k < 0 || p->top_k > kMaxTopKNet)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "top_k out of [0, 8] (0 = all views)");
  if (p->iterations < 0 || p->propagation_reach < 1 || p->propagation_reach > 64)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad iterations / propagation_reach");
  if (!(p->value_scale > 0.0)) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "value_scale must be > 0");
  if (p->tile_schedule < -1 || p->tile_schedule >= kNumSchedules)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "tile_schedule out of [-1, 2]");
  if (p->propagation_reach > kMaxReach)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "propagation_reach above 16");
  if (p->occupancy_cap < -1 || p->occupancy_cap > 64)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "occupancy_cap out of [-1, 64]");
  if (p->view_phased_iterations < -1)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "view_phased_iterations below -1");
  if (p->binned_phases < -1)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "binned_phases below -1");
  return RECON_MVS_OK;
}

// Builds the kernel arguments for ref view `ref` and its sources.
int32_t MakeArgs(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs,
                 const recon_mvs_plane_params* p, PlaneArgs* a, const double* pose_delta = nullptr,
                 PairGeometry* geo = nullptr) {
  int32_t st = CheckParams(ctx, p);
  if (st != RECON_MVS_OK) return st;
  const int nv = static_cast<int>(ctx->views.size());
  if (ref < 0 || ref >= nv) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "ref_view out of range");
  if (nsrc < 1 || nsrc > kMaxViews || !srcs)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "num_src out of [1, 16]");
  for (int k = 0; k < nsrc; ++k)
    if (srcs[k] < 0 || srcs[k] >= nv || srcs[k] == ref)
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad source view index");
  const ResidentView& R = *ctx->views[ref];
  std::memset(a, 0, sizeof(*a));
  a->w = R.w;
  a->h = R.h;
  a->ref_pitch = R.pitch;
  a->ref_img = R.pixels;
  a->fx = R.cam.fx; a->fy = R.cam.fy; a->cx = R.cam.cx; a->cy = R.cam.cy;
  a->kinv[0] = 1.0 / R.cam.fx; a->kinv[1] = 1.0 / R.cam.fy;
  a->kinv[2] = -R.cam.cx / R.cam.fx; a->kinv[3] = -R.cam.cy / R.cam.fy;
  a->ffx = static_cast<float>(R.cam.fx); a->ffy = static_cast<float>(R.cam.fy);
  a->fcx = static_cast<float>(R.cam.cx); a->fcy = static_cast<float>(R.cam.cy);
  a->nsrc = nsrc;
  a->top_k = p->top_k == 0 ? nsrc : std::min(p->top_k, nsrc);
  // K_r^-1
  const double Ki[9] = {1.0 / R.cam.fx, 0.0, -R.cam.cx / R.cam.fx,
                        0.0, 1.0 / R.cam.fy, -R.cam.cy / R.cam.fy,
                        0.0, 0.0, 1.0};
  for (int k = 0; k < nsrc; ++k) {
    const ResidentView& S = *ctx->views[srcs[k]];
    // source pose, perturbed on the left when asked: R' = E R, t' = E t + tau
    double SR[9], St[3];
    std::memcpy(SR, S.cam.R, sizeof(SR));
    std::memcpy(St, S.cam.t, sizeof(St));
    if (pose_delta) {
      const double* d = pose_delta + 6 * k;
      double E[9];
      So3Exp(d, E);
      for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j)
          SR[3 * i + j] = E[3 * i] * S.cam.R[j] + E[3 * i + 1] * S.cam.R[3 + j] + E[3 * i + 2] * S.cam.R[6 + j];
        St[i] = E[3 * i] * S.cam.t[0] + E[3 * i + 1] * S.cam.t[1] + E[3 * i + 2] * S.cam.t[2] + d[3 + i];
      }
    }
    a->src_img[k] = S.pixels;
    a->src_pitch[k] = S.pitch;
    a->src_w[k] = S.w;
    a->src_h[k] = S.h;
    a->src_lim[k][0] = S.w - 1.0;
    a->src_lim[k][1] = S.h - 1.0;
    a->src_lim[k][2] = S.w - 1.0 - 1e-4;
    a->src_lim[k][3] = S.h - 1.0 - 1e-4;
    // R_rel = R_s R_r^T, t_rel = t_s - R_rel t_r (FP64 composition per view pair)
    double Rrel[9], trel[3];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int l = 0; l < 3; ++l) s += SR[3 * i + l] * R.cam.R[3 * j + l];
        Rrel[3 * i + j] = s;
      }
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l) s += Rrel[3 * i + l] * R.cam.t[l];
      trel[i] = St[i] - s;
    }
    if (geo) {
      std::memcpy(geo->Rrel[k], Rrel, sizeof(Rrel));
      std::memcpy(geo->trel[k], trel, sizeof(trel));
    }
    const double Ks[9] = {S.cam.fx, 0.0, S.cam.cx, 0.0, S.cam.fy, S.cam.cy, 0.0, 0.0, 1.0};
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
        a->A[k][3 * i + j] = s;
      }
    for (int i = 0; i < 3; ++i) {
      double s = 0.0;
      for (int l = 0; l < 3; ++l) s += Ks[3 * i + l] * trel[l];
      a->b[k][i] = s;
    }
  }
  a->min_depth = static_cast<float>(p->min_depth_m);
  a->max_depth = static_cast<float>(p->max_depth_m);
  a->inv_min = static_cast<float>(1.0 / p->max_depth_m);
  a->inv_max = static_cast<float>(1.0 / p->min_depth_m);
  const double eps = 1e-9 * (p->value_scale / 255.0) * (p->value_scale / 255.0);
  a->eps2 = static_cast<float>(eps * eps);
  a->photo_reject = static_cast<float>(3.0 * p->photo_threshold);
  a->seed = p->seed;
  a->reach = p->propagation_reach;
  a->depth_perturb = static_cast<float>(p->depth_perturb);
  a->normal_perturb = static_cast<float>(p->normal_perturb);
  return RECON_MVS_OK;
}

int32_t EnsureState(recon_mvs_hip_context* ctx, int view) {
  if (static_cast<int>(ctx->states.size()) < static_cast<int>(ctx->views.size()))
    ctx->states.resize(ctx->views.size(), nullptr);
  PlaneState*& s = ctx->states[view];
  const ResidentView& V = *ctx->views[view];
  if (s && s->w == V.w && s->h == V.h) return RECON_MVS_OK;
  FreeState(s);
  s = new PlaneState();
  s->w = V.w;
  s->h = V.h;
  const size_t P = static_cast<size_t>(V.w) * V.h;
  if (hipMalloc(&s->plane, P * sizeof(float4)) != hipSuccess ||
      hipMalloc(&s->prev, 4 * 2 * static_cast<size_t>(V.h) * ((V.w + 1) / 2) * sizeof(float4)) != hipSuccess ||
      hipMalloc(&s->cand, 4 * static_cast<size_t>(V.h) * ((V.w + 1) / 2) * sizeof(float4)) != hipSuccess ||
      hipMalloc(&s->cost, P * sizeof(float)) != hipSuccess ||
      hipMalloc(&s->support, P) != hipSuccess || hipMalloc(&s->geo_support, P) != hipSuccess) {
    FreeState(s);
    s = nullptr;
    return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "plane state allocation failed");
  }
  return RECON_MVS_OK;
}

enum KernelId {
  kSweep = 0, kInit = 1, kRefine = 2, kGeo = 3, kScore = 4, kEval = 5, kBatch = 6,
  // view-phased half-sweep parts (timed together as kSweep)
  kPhasePrep = 8, kPhaseView = 9, kPhaseFinal = 10
};

// Arguments of one persistent batched launch.
struct BatchLaunch {
  const PlaneArgs* args;
  const int2* units;
  const int* qoff;
  unsigned int* counters;
  int phase;
  const int32_t* cancel;
};

// Arguments of one part of a view-phased half-sweep.
struct PhaseLaunch {
  PhaseState st;
  const int* order;  // kPhaseView: this view's tile order
  int view, color, iter;
};

// Tile geometry of schedule `s` (see kNumSchedules).
inline void ScheduleTile(int s, int* tw, int* th) {
  *tw = s == 1 ? 32 : (s == 2 ? 128 : 64);
  *th = 512 / *tw;
}

// Dynamic LDS that limits a kernel with `stat` bytes of static LDS to `cap`
// resident workgroups per CU (0: no limit).  The padding is never touched; it
// only makes the (cap + 1)-th workgroup not fit in the CU's LDS.
inline size_t CapBytes(int cap, size_t stat) {
  static int lds_cu = -1, lds_block = 0;
  if (lds_cu < 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess)
      lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      lds_block = 0;
  }
  if (cap <= 0 || lds_cu <= 0) return 0;
  const size_t per_wg = (static_cast<size_t>(lds_cu) / cap) & ~static_cast<size_t>(1023);
  const size_t want = per_wg > stat + 1024 ? per_wg - 1024 - stat : 0;
  const size_t room = static_cast<size_t>(lds_block) > stat ? static_cast<size_t>(lds_block) - stat : 0;
  return std::min(want, room);
}

// Phase configuration of a launch: occupancy cap (low byte; 0 = none) plus
// kBinnedCfg for the binned form of init / half-sweep (11x11 stride-2 window;
// other windows run the fused form, which computes the same results).
constexpr int kBinnedCfg = 1 << 8;
inline int CfgCap(int cfg) { return cfg & 0xFF; }
inline bool CfgBinned(int cfg) { return (cfg & kBinnedCfg) != 0; }
// Whether the binned form exists for p's window.
inline bool HasBinned(const recon_mvs_plane_params* p) { return p->patch_size == 11 && p->window_stride == 2; }

// Workgroups of a persistent launch: as many as fit on the device at once
// (with `cap` workgroups per CU at most).
template <int HALF, int STRIDE, int FMT, int TW, int TH, bool BINNED>
int BatchGrid(int cap) {
  static int nb = -1, cus = 1;
  if (nb < 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_plane_batch<HALF, STRIDE, FMT, TW, TH, BINNED>, 256,
                                                     0) != hipSuccess)
      nb = 1;
  }
  const int per_cu = std::max(1, cap > 0 ? std::min(cap, nb) : nb);
  return per_cu * std::max(1, cus);
}

// Static LDS of a kernel as compiled (the compiler may place private arrays
// in LDS too); on failure a size that disables the padding.
inline size_t StaticLds(const void* kernel) {
  hipFuncAttributes at;
  if (hipFuncGetAttributes(&at, kernel) != hipSuccess) return 1u << 30;
  return at.sharedSizeBytes;
}

template <int HALF, int STRIDE, int FMT, int TW, int TH>
void LaunchTile(KernelId id, PlaneArgs a, hipStream_t s, float f0, int i0, int n, const float* hyps,
                float* cost, int* nvalid, float* view_ncc, const BatchLaunch* bl, int cfg,
                const PhaseLaunch* pl) {
  constexpr bool kHasBinned = (HALF == 5 && STRIDE == 2);
  const bool binned = kHasBinned && CfgBinned(cfg) && (id == kSweep || id == kInit || id == kBatch);
  const int cap = CfgCap(cfg);
  // padding for the occupancy cap, from each kernel's own static LDS
  size_t pad = 0;
  if (cap > 0) {
    static const size_t lds_sweep = StaticLds(reinterpret_cast<const void*>(k_plane_sweep<HALF, STRIDE, FMT, TW, TH>));
    static const size_t lds_init = StaticLds(reinterpret_cast<const void*>(k_plane_init<HALF, STRIDE, FMT, TW, TH>));
    static const size_t lds_batch = StaticLds(reinterpret_cast<const void*>(k_plane_batch<HALF, STRIDE, FMT, TW, TH, false>));
    if (binned) {
      if constexpr (kHasBinned) {
        static const size_t lds_bsweep = StaticLds(reinterpret_cast<const void*>(k_plane_sweep_binned<HALF, STRIDE, FMT, TW, TH>));
        static const size_t lds_binit = StaticLds(reinterpret_cast<const void*>(k_plane_init_binned<HALF, STRIDE, FMT, TW, TH>));
        static const size_t lds_bbatch = StaticLds(reinterpret_cast<const void*>(k_plane_batch<HALF, STRIDE, FMT, TW, TH, true>));
        pad = CapBytes(cap, id == kSweep ? lds_bsweep : id == kInit ? lds_binit : lds_bbatch);
      }
    } else if (id == kSweep) pad = CapBytes(cap, lds_sweep);
    else if (id == kInit) pad = CapBytes(cap, lds_init);
    else if (id == kBatch) pad = CapBytes(cap, lds_batch);
  }
  if (id == kBatch) {
    hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, bl->counters);
    if constexpr (kHasBinned) {
      if (binned) {
        hipLaunchKernelGGL((k_plane_batch<HALF, STRIDE, FMT, TW, TH, true>),
                           dim3(BatchGrid<HALF, STRIDE, FMT, TW, TH, true>(cap)), dim3(256), pad, s, bl->args,
                           bl->units, bl->qoff, bl->counters, bl->phase, bl->cancel);
        return;
      }
    }
    hipLaunchKernelGGL((k_plane_batch<HALF, STRIDE, FMT, TW, TH, false>),
                       dim3(BatchGrid<HALF, STRIDE, FMT, TW, TH, false>(cap)), dim3(256), pad, s, bl->args,
                       bl->units, bl->qoff, bl->counters, bl->phase, bl->cancel);
    return;
  }
  a.tiles_x = (a.w + TW - 1) / TW;
  a.tiles_y = (a.h + TH - 1) / TH;
  a.num_tiles = a.tiles_x * a.tiles_y;
  a.tiles_per_xcd = (a.num_tiles + 7) / 8;
  const dim3 grid(a.tiles_per_xcd * 8), block(256);
  if constexpr (kHasBinned) {
    if (binned) {
      if (id == kSweep) hipLaunchKernelGGL((k_plane_sweep_binned<HALF, STRIDE, FMT, TW, TH>), grid, block, pad, s, a);
      else hipLaunchKernelGGL((k_plane_init_binned<HALF, STRIDE, FMT, TW, TH>), grid, block, pad, s, a);
      return;
    }
  }
  switch (id) {
    case kSweep: hipLaunchKernelGGL((k_plane_sweep<HALF, STRIDE, FMT, TW, TH>), grid, block, pad, s, a); break;
    case kInit: hipLaunchKernelGGL((k_plane_init<HALF, STRIDE, FMT, TW, TH>), grid, block, pad, s, a); break;
    case kScore: hipLaunchKernelGGL((k_plane_score<HALF, STRIDE, FMT, TW, TH>), grid, block, 0, s, a); break;
    case kRefine:
      hipLaunchKernelGGL((k_plane_refine<HALF, STRIDE, FMT, TW, TH>), grid, block, 0, s, a, f0, i0);
      break;
    case kEval:
      hipLaunchKernelGGL((k_plane_eval<HALF, STRIDE, FMT>), dim3((n + 63) / 64), dim3(64), 0, s, a,
                         n, hyps, cost, nvalid, view_ncc);
      break;
    case kPhasePrep:
      if constexpr (HALF == 5 && STRIDE == 2)
        hipLaunchKernelGGL((k_phase_prep<HALF, STRIDE, FMT, TW, TH>), grid, block, 0, s, a, pl->st, pl->color,
                           pl->iter);
      break;
    case kPhaseView:
      if constexpr (HALF == 5 && STRIDE == 2)
        hipLaunchKernelGGL((k_phase_view<HALF, STRIDE, FMT, TW, TH>), grid, block, pad, s, a, pl->st, pl->order,
                           pl->view, pl->color);
      break;
    case kPhaseFinal:
      if constexpr (HALF == 5 && STRIDE == 2)
        hipLaunchKernelGGL((k_phase_final<HALF, STRIDE, FMT, TW, TH>), grid, block, 0, s, a, pl->st, pl->color);
      break;
    default: break;
  }
}

template <int HALF, int STRIDE, int FMT>
bool LaunchT(int schedule, KernelId id, const PlaneArgs& a, hipStream_t s, float f0, int i0, int n,
             const float* hyps, float* cost, int* nvalid, float* view_ncc, const BatchLaunch* bl,
             int cfg, const PhaseLaunch* pl) {
  // the 11x11 stride-2 window is compiled for every tile schedule, the others
  // for the default one
  constexpr bool kAllSchedules = (HALF == 5 && STRIDE == 2);
  switch (schedule) {
    case 0: LaunchTile<HALF, STRIDE, FMT, 64, 8>(id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl); return true;
    case 1:
      if constexpr (kAllSchedules) { LaunchTile<HALF, STRIDE, FMT, 32, 16>(id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl); return true; }
      return false;
    case 2:
      if constexpr (kAllSchedules) { LaunchTile<HALF, STRIDE, FMT, 128, 4>(id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl); return true; }
      return false;
    default: return false;
  }
}

template <int HALF, int STRIDE>
bool LaunchFmt(int fmt, int schedule, KernelId id, const PlaneArgs& a, hipStream_t s, float f0, int i0,
               int n, const float* hyps, float* cost, int* nvalid, float* view_ncc, const BatchLaunch* bl,
               int cfg, const PhaseLaunch* pl) {
  // FP16 / BF16 storage is compiled for the 11x11 stride-2 and 7x7 windows,
  // 8-bit storage for the 11x11 stride-2 window
  constexpr bool kAllFormats = (HALF == 5 && STRIDE == 2) || (HALF == 3 && STRIDE == 1);
  switch (fmt) {
    case RECON_MVS_PIXEL_F32: return LaunchT<HALF, STRIDE, RECON_MVS_PIXEL_F32>(schedule, id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
    case RECON_MVS_PIXEL_F16:
      if constexpr (kAllFormats) return LaunchT<HALF, STRIDE, RECON_MVS_PIXEL_F16>(schedule, id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
      return false;
    case RECON_MVS_PIXEL_BF16:
      if constexpr (kAllFormats) return LaunchT<HALF, STRIDE, RECON_MVS_PIXEL_BF16>(schedule, id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
      return false;
    case RECON_MVS_PIXEL_U8:  // 8-bit storage: the 11x11 stride-2 window
      if constexpr (HALF == 5 && STRIDE == 2) return LaunchT<HALF, STRIDE, RECON_MVS_PIXEL_U8>(schedule, id, a, s, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
      return false;
    default: return false;
  }
}

// Window shapes compiled in: (patch, stride) in {(11,2), (11,1), (7,1), (5,1), (7,2), (9,2)}.
// cfg: phase configuration (occupancy cap, binned form) of init / sweep /
// batch launches.
int32_t Launch(recon_mvs_hip_context* ctx, const recon_mvs_plane_params* p, KernelId id,
               const PlaneArgs& a, float f0 = 0.0f, int i0 = 0, int n = 0,
               const float* hyps = nullptr, float* cost = nullptr, int* nvalid = nullptr,
               float* view_ncc = nullptr, const BatchLaunch* bl = nullptr, int schedule = -1,
               int cfg = 0, const PhaseLaunch* pl = nullptr) {
  const int fmt = ctx->storage_format;
  const int sch = schedule >= 0 ? schedule : p->tile_schedule;
  hipEvent_t eb, ee;
  TimerBegin(ctx, static_cast<int>(id), &eb, &ee);
  bool ok = false;
  const int ps = p->patch_size, st = p->window_stride;
  if (ps == 11 && st == 2) ok = LaunchFmt<5, 2>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
#ifndef RMVS_BENCH_WINDOW_ONLY  // (development builds compile the 11x11 stride-2 window only)
  else if (ps == 11 && st == 1) ok = LaunchFmt<5, 1>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
  else if (ps == 7 && st == 1) ok = LaunchFmt<3, 1>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
  else if (ps == 5 && st == 1) ok = LaunchFmt<2, 1>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
  else if (ps == 7 && st == 2) ok = LaunchFmt<3, 2>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
  else if (ps == 9 && st == 2) ok = LaunchFmt<4, 2>(fmt, sch, id, a, ctx->stream, f0, i0, n, hyps, cost, nvalid, view_ncc, bl, cfg, pl);
#endif
  if (!ok)
    return Fail(ctx, RECON_MVS_ERR_NOT_IMPLEMENTED,
                "window (patch_size, window_stride) / storage format / tile schedule not compiled in");
  TimerEnd(ctx, static_cast<int>(id), eb, ee);
  RMVS_LAUNCH_CHECK(ctx);
  return RECON_MVS_OK;
}

// Direction of the epipolar lines of source S in reference R at the image
// centre (unit vector; false when the epipole is at the centre).  The lines
// run through the epipole e = K_r (R_r C_s + t_r); the direction is defined
// also when e is at infinity.
bool EpipolarDirection(const ResidentView& R, const ResidentView& S, double* ux, double* uy) {
  double C[3], P[3];
  for (int i = 0; i < 3; ++i)  // source centre C = -R_s^T t_s
    C[i] = -(S.cam.R[i] * S.cam.t[0] + S.cam.R[3 + i] * S.cam.t[1] + S.cam.R[6 + i] * S.cam.t[2]);
  for (int i = 0; i < 3; ++i)
    P[i] = R.cam.R[3 * i] * C[0] + R.cam.R[3 * i + 1] * C[1] + R.cam.R[3 * i + 2] * C[2] + R.cam.t[i];
  const double ex = R.cam.fx * P[0] + R.cam.cx * P[2], ey = R.cam.fy * P[1] + R.cam.cy * P[2];
  const double dx = ex - 0.5 * R.w * P[2], dy = ey - 0.5 * R.h * P[2];
  const double len = std::hypot(dx, dy);
  if (!(len > 1e-12)) return false;
  *ux = dx / len;
  *uy = dy / len;
  return true;
}

// Tile schedule of a view-phased launch: tiles long along the epipolar lines
// (128x4 for lines near the rows, 32x16 near the columns, 64x8 between), so a
// tile's unconverged hypotheses sample a narrow band of the source.
int PhaseSchedule(const ResidentView& R, const ResidentView& S) {
  double ux = 1.0, uy = 0.0;
  if (!EpipolarDirection(R, S, &ux, &uy)) return 0;
  if (std::fabs(ux) > 2.0 * std::fabs(uy)) return 2;
  if (std::fabs(uy) > 2.0 * std::fabs(ux)) return 1;
  return 0;
}

// Tile order of a view-phased launch (ty << 16 | tx per entry): strips of
// tiles along the epipolar lines, taken strip by strip.  Dealt to the XCDs in
// runs (k_phase_view), each XCD's tiles in flight then form a stretch of a
// strip whose source footprint is one band of the source image.
std::vector<int> PhaseTileOrder(const ResidentView& R, const ResidentView& S, int tw, int th) {
  const int tiles_x = (R.w + tw - 1) / tw, tiles_y = (R.h + th - 1) / th;
  std::vector<int> order;
  order.reserve(static_cast<size_t>(tiles_x) * tiles_y);
  double ux, uy;
  if (!EpipolarDirection(R, S, &ux, &uy)) {
    // epipole at the image centre (motion along the axis): down the columns
    for (int t = 0; t < tiles_x * tiles_y; ++t) order.push_back(((t % tiles_y) << 16) | (t / tiles_y));
    return order;
  }
  const double nx = -uy, ny = ux, mx = 0.5 * R.w, my = 0.5 * R.h;
  const double thick = std::fabs(nx) * tw + std::fabs(ny) * th;  // a tile's extent across the lines
  std::vector<std::pair<std::pair<long long, double>, int>> keyed;
  keyed.reserve(order.capacity());
  for (int ty = 0; ty < tiles_y; ++ty)
    for (int tx = 0; tx < tiles_x; ++tx) {
      const double px = (tx + 0.5) * tw - mx, py = (ty + 0.5) * th - my;
      const long long strip = static_cast<long long>(std::floor((nx * px + ny * py) / thick));
      keyed.push_back({{strip, ux * px + uy * py}, (ty << 16) | tx});
    }
  std::sort(keyed.begin(), keyed.end());
  for (const auto& k : keyed) order.push_back(k.second);
  return order;
}

// One view-phased half-sweep (k_phase_prep, one k_phase_view per source,
// k_phase_final) of the view `a` describes; timed as one kSweep.
int32_t PhasedHalfSweep(recon_mvs_hip_context* ctx, const recon_mvs_plane_params* p, PlaneArgs a, int ref,
                        const int32_t* srcs, int sched, int iter, int color) {
  std::string key = std::to_string(ref) + "/" + std::to_string(a.w) + "x" + std::to_string(a.h) + ":";
  for (int v = 0; v < a.nsrc; ++v) key += std::to_string(srcs[v]) + ",";
  int*& orders = ctx->phase_orders[key];
  std::vector<int> vsched(a.nsrc), voff(a.nsrc);
  {
    std::vector<int> host;
    for (int v = 0; v < a.nsrc; ++v) {
      vsched[v] = PhaseSchedule(*ctx->views[ref], *ctx->views[srcs[v]]);
      int tw, th;
      ScheduleTile(vsched[v], &tw, &th);
      voff[v] = static_cast<int>(host.size());
      if (!orders) {
        const std::vector<int> o = PhaseTileOrder(*ctx->views[ref], *ctx->views[srcs[v]], tw, th);
        host.insert(host.end(), o.begin(), o.end());
      } else {
        host.resize(host.size() + static_cast<size_t>((a.w + tw - 1) / tw) * ((a.h + th - 1) / th));
      }
    }
    if (!orders) {
      if (hipMalloc(&orders, host.size() * sizeof(int)) != hipSuccess) {
        orders = nullptr;
        return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "phase order allocation failed");
      }
      RMVS_HIP(ctx, hipMemcpy(orders, host.data(), host.size() * sizeof(int), hipMemcpyHostToDevice));
    }
  }
  PhaseLaunch pl{};
  PhaseState& st = pl.st;
  st.cw = (a.w + 1) / 2;
  st.n = a.h * st.cw;
  st.K = a.top_k < a.nsrc ? a.top_k : 0;
  const size_t n8 = static_cast<size_t>(st.n) * 8;
  char* base = static_cast<char*>(Ensure(ctx, kSlotPhase, n8 * (sizeof(float4) + sizeof(float) * (st.K + 1) + sizeof(int))));
  if (!base) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "phase state allocation failed");
  st.cand = reinterpret_cast<float4*>(base);
  st.topk = reinterpret_cast<float*>(base + n8 * sizeof(float4));
  st.csum = st.topk + n8 * st.K;
  st.nvalid = reinterpret_cast<int*>(st.csum + n8);
  pl.color = color;
  pl.iter = iter;
  a.color = color;
  a.iter = iter;
  hipEvent_t eb, ee;
  TimerBegin(ctx, kSweep, &eb, &ee);
  int32_t r = Launch(ctx, p, kPhasePrep, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched, 0, &pl);
  for (int v = 0; v < a.nsrc && r == RECON_MVS_OK; ++v) {
    pl.view = v;
    pl.order = orders + voff[v];
    r = Launch(ctx, p, kPhaseView, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, vsched[v], 0, &pl);
  }
  if (r == RECON_MVS_OK)
    r = Launch(ctx, p, kPhaseFinal, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched, 0, &pl);
  TimerEnd(ctx, kSweep, eb, ee);
  return r;
}

double MeanBaselineViews(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs) {
  std::vector<recon_mvs_camera> c(nsrc);
  for (int k = 0; k < nsrc; ++k) c[k] = ctx->views[srcs[k]]->cam;
  return MeanBaseline(ctx->views[ref]->cam, c.data(), nsrc);
}

std::string TuneKey(const recon_mvs_hip_context* ctx, const PlaneArgs& a, const recon_mvs_plane_params* p) {
  char buf[128];
  std::snprintf(buf, sizeof(buf), "%dx%d/%d/%d/%d/%d/%d", a.w, a.h, a.nsrc, a.top_k, p->patch_size,
                p->window_stride, ctx->storage_format);
  return buf;
}

// Times random init + two half-sweeps of view `a` under every compiled tile
// schedule and returns the fastest (the state of the slot is overwritten).
// The phases run binned where that form may be used, as the run will.
int32_t Autotune(recon_mvs_hip_context* ctx, const PlaneArgs& a, const recon_mvs_plane_params* p,
                 int* best, double* ms_out) {
  const int cfg = HasBinned(p) && p->binned_phases != 0 ? kBinnedCfg : 0;
  hipEvent_t e0, e1;
  RMVS_HIP(ctx, hipEventCreate(&e0));
  RMVS_HIP(ctx, hipEventCreate(&e1));
  const bool timing = ctx->timing;
  ctx->timing = false;
  double best_ms = 1e300;
  *best = 0;
  int32_t st = RECON_MVS_OK;
  for (int s = 0; s < kNumSchedules; ++s) {
    if (ms_out) ms_out[s] = -1.0;
    // warm-up launch (code load), then the timed one
    PlaneArgs b = a;
    if (Launch(ctx, p, kInit, b, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, s, cfg) !=
        RECON_MVS_OK) {
      ctx->error.clear();
      continue;
    }
    (void)hipEventRecord(e0, ctx->stream);
    for (int c = 0; c < 2 && st == RECON_MVS_OK; ++c) {
      b.iter = 0;
      b.color = c;
      st = Launch(ctx, p, kSweep, b, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, s, cfg);
    }
    (void)hipEventRecord(e1, ctx->stream);
    if (st != RECON_MVS_OK) break;
    float ms = 0.0f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "autotune timing failed");
      break;
    }
    if (ms_out) ms_out[s] = ms;
    if (ms < best_ms) {
      best_ms = ms;
      *best = s;
    }
  }
  ctx->timing = timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

// The schedule to use for view `a`: p->tile_schedule, or the autotuned one
// (cached per size / views / window / format) when it is -1.
int32_t ResolveSchedule(recon_mvs_hip_context* ctx, const PlaneArgs& a,
                        const recon_mvs_plane_params* p, int* sched) {
  if (p->tile_schedule >= 0) {
    *sched = p->tile_schedule;
    return RECON_MVS_OK;
  }
  const std::string key = TuneKey(ctx, a, p);
  auto it = ctx->tuned_schedule.find(key);
  if (it != ctx->tuned_schedule.end()) {
    *sched = it->second;
    return RECON_MVS_OK;
  }
  int best = 0;
  const int32_t st = Autotune(ctx, a, p, &best, nullptr);
  if (st != RECON_MVS_OK) return st;
  ctx->tuned_schedule[key] = best;
  *sched = best;
  return RECON_MVS_OK;
}

// Times one full PatchMatch run of view `a` (init + every half-sweep) with
// phase configuration `cap` (occupancy cap, binned form) in every phase;
// ms[phase] receives each launch's time.
int32_t TimePhases(recon_mvs_hip_context* ctx, PlaneArgs a, const recon_mvs_plane_params* p,
                   int sched, int cap, std::vector<float>* ms) {
  const int nph = 1 + 2 * p->iterations;
  std::vector<hipEvent_t> ev(nph + 1, nullptr);
  int32_t st = RECON_MVS_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event create failed");
  if (st == RECON_MVS_OK && hipEventRecord(ev[0], ctx->stream) != hipSuccess)
    st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event record failed");
  for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph) {
    if (ph == 0) {
      st = Launch(ctx, p, kInit, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched, cap);
    } else {
      a.iter = (ph - 1) / 2;
      a.color = (ph - 1) % 2;
      st = Launch(ctx, p, kSweep, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched, cap);
    }
    if (st == RECON_MVS_OK && hipEventRecord(ev[ph + 1], ctx->stream) != hipSuccess)
      st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event record failed");
  }
  if (st == RECON_MVS_OK && hipEventSynchronize(ev[nph]) != hipSuccess)
    st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "occupancy tuning run failed");
  ms->assign(nph, 0.0f);
  for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph)
    if (hipEventElapsedTime(&(*ms)[ph], ev[ph], ev[ph + 1]) != hipSuccess)
      st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event timing failed");
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return st;
}

// Configuration (occupancy cap, binned form) of every phase of a PatchMatch
// run (index 0: init; 1 + 2 it + c: half-sweep c of iteration it), as
// recon_mvs_plane_params.occupancy_cap and binned_phases describe.  Tuning
// runs the view under caps 0 (none), 1 and 2, fused and (binned_phases = -1)
// binned, and keeps each phase's fastest; the phases compute the same
// results under any configuration.
int32_t ResolveCaps(recon_mvs_hip_context* ctx, const PlaneArgs& a, const recon_mvs_plane_params* p,
                    int sched, std::vector<int>* caps) {
  const int nph = 1 + 2 * p->iterations;
  caps->assign(nph, 0);
  const bool binnable = HasBinned(p);
  // forced binned phases keep the binned form whatever the caps say
  const int forced = binnable && p->binned_phases > 0 ? std::min(nph, p->binned_phases) : 0;
  auto force = [&]() {
    for (int ph = 0; ph < forced; ++ph) (*caps)[ph] = kBinnedCfg;
  };
  if (p->occupancy_cap < 0) {
    force();
    return RECON_MVS_OK;
  }
  const int ci = p->capped_iterations < 0 ? 1 : p->capped_iterations;
  const int ncapped = std::min(nph, 1 + 2 * ci);
  if (p->occupancy_cap > 0 || p->tile_schedule >= 0) {
    const int cap = p->occupancy_cap > 0 ? p->occupancy_cap
                                         : (ctx->storage_format == RECON_MVS_PIXEL_F32 ? 1 : 0);
    for (int ph = 0; ph < ncapped; ++ph) (*caps)[ph] = cap;
    force();
    return RECON_MVS_OK;
  }
  const std::string key = TuneKey(ctx, a, p) + "/" + std::to_string(sched) + "/" +
                          std::to_string(p->iterations) + "/" + std::to_string(p->binned_phases);
  auto it = ctx->tuned_caps.find(key);
  if (it != ctx->tuned_caps.end()) {
    *caps = it->second;
    return RECON_MVS_OK;
  }
  const bool timing = ctx->timing;
  ctx->timing = false;
  std::vector<int> trials = {0, 1, 2};
  if (binnable && p->binned_phases < 0)
    for (int c : {kBinnedCfg, kBinnedCfg | 1, kBinnedCfg | 2}) trials.push_back(c);
  std::vector<float> best(nph, 3.0e38f), ms;
  int32_t st = TimePhases(ctx, a, p, sched, 0, &ms);  // warm-up (code load)
  for (size_t t = 0; t < trials.size() && st == RECON_MVS_OK; ++t) {
    st = TimePhases(ctx, a, p, sched, trials[t], &ms);
    for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph)
      if (ms[ph] < best[ph]) {
        best[ph] = ms[ph];
        (*caps)[ph] = trials[t];
      }
  }
  ctx->timing = timing;
  if (st != RECON_MVS_OK) return st;
  force();
  ctx->tuned_caps[key] = *caps;
  return RECON_MVS_OK;
}

// Work units of a batch under tile schedule `sched` (sets each job's tile
// counts): one queue per XCD holding a vertical band of every job, jobs in
// `order` -- tile c in column order goes to queue c / tiles_per_xcd
// (TileOf's layout for a single view).
void BatchUnits(std::vector<PlaneArgs>* args, const std::vector<int>& order, int sched,
                std::vector<int2>* units, std::vector<int>* qoff) {
  int tw, th;
  ScheduleTile(sched, &tw, &th);
  std::vector<std::vector<int2>> queue(kXcds);
  for (int j : order) {
    PlaneArgs& a = (*args)[j];
    a.tiles_x = (a.w + tw - 1) / tw;
    a.tiles_y = (a.h + th - 1) / th;
    a.num_tiles = a.tiles_x * a.tiles_y;
    a.tiles_per_xcd = (a.num_tiles + kXcds - 1) / kXcds;
    for (int c = 0; c < a.num_tiles; ++c) {
      const int tx = c / a.tiles_y, ty = c % a.tiles_y;
      queue[c / a.tiles_per_xcd].push_back(make_int2(j, ty * a.tiles_x + tx));
    }
  }
  units->clear();
  qoff->assign(kXcds + 1, 0);
  for (int q = 0; q < kXcds; ++q) {
    units->insert(units->end(), queue[q].begin(), queue[q].end());
    (*qoff)[q + 1] = static_cast<int>(units->size());
  }
}

// Times every phase of a batch (init, then each half-sweep) as eager
// launches with events between them, phase ph capped at caps[ph] workgroups
// per CU (0: none); ms[ph] receives each launch's time.
int32_t TimeBatchPhases(recon_mvs_hip_context* ctx, const recon_mvs_plane_params* p, const PlaneArgs& a0,
                        BatchLaunch bl, int sched, const std::vector<int>& caps, std::vector<float>* ms) {
  const int nph = 1 + 2 * p->iterations;
  std::vector<hipEvent_t> ev(nph + 1, nullptr);
  int32_t st = RECON_MVS_OK;
  for (auto& e : ev)
    if (hipEventCreate(&e) != hipSuccess) st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event create failed");
  if (st == RECON_MVS_OK && hipEventRecord(ev[0], ctx->stream) != hipSuccess)
    st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event record failed");
  for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph) {
    bl.phase = ph - 1;
    st = Launch(ctx, p, kBatch, a0, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, &bl, sched, caps[ph]);
    if (st == RECON_MVS_OK && hipEventRecord(ev[ph + 1], ctx->stream) != hipSuccess)
      st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event record failed");
  }
  if (st == RECON_MVS_OK && hipEventSynchronize(ev[nph]) != hipSuccess)
    st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "batch tuning run failed");
  ms->assign(nph, 0.0f);
  for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph)
    if (hipEventElapsedTime(&(*ms)[ph], ev[ph], ev[ph + 1]) != hipSuccess)
      st = Fail(ctx, RECON_MVS_ERR_GPU_FAILURE, "event timing failed");
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& e : ev)
    if (e) (void)hipEventDestroy(e);
  return st;
}

}  // namespace

namespace rmvs {

void So3Exp(const double w[3], double E[9]) {
  const double th = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  const double K[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
  double K2[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      K2[3 * i + j] = K[3 * i] * K[j] + K[3 * i + 1] * K[3 + j] + K[3 * i + 2] * K[6 + j];
  // sin(th)/th and (1 - cos(th))/th^2, by their series below 1e-6 rad
  const double s = th < 1e-6 ? 1.0 - th * th / 6.0 : std::sin(th) / th;
  const double c = th < 1e-6 ? 0.5 - th * th / 24.0 : (1.0 - std::cos(th)) / (th * th);
  for (int i = 0; i < 9; ++i) E[i] = (i % 4 == 0 ? 1.0 : 0.0) + s * K[i] + c * K2[i];
}

int32_t PlaneMakeArgs(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs,
                      const recon_mvs_plane_params* p, PlaneArgs* a, const double* pose_delta,
                      PairGeometry* geo) {
  return MakeArgs(ctx, ref, nsrc, srcs, p, a, pose_delta, geo);
}

int32_t PlaneEnsureState(recon_mvs_hip_context* ctx, int view) { return EnsureState(ctx, view); }

double PlaneMeanBaseline(recon_mvs_hip_context* ctx, int ref, int nsrc, const int32_t* srcs) {
  return MeanBaselineViews(ctx, ref, nsrc, srcs);
}

}  // namespace rmvs

struct recon_mvs_hip_batch {
  recon_mvs_plane_params params{};
  int schedule = 0;
  std::vector<int> refs;
  PlaneArgs* d_args = nullptr;
  int2* d_units = nullptr;
  int num_units = 0;
  int* d_qoff = nullptr;             // kXcds + 1 queue offsets into d_units
  unsigned int* d_counter = nullptr; // kXcds queue counters, kCounterStride apart
  hipGraphExec_t exec = nullptr;
};

void DestroyPlaneState(recon_mvs_hip_context* ctx) {
  if (!ctx->phase_orders.empty()) {
    (void)hipStreamSynchronize(ctx->stream);
    for (auto& kv : ctx->phase_orders)
      if (kv.second) (void)hipFree(kv.second);
    ctx->phase_orders.clear();
  }
  for (auto* s : ctx->states) FreeState(s);
  ctx->states.clear();
  for (auto* v : ctx->views) FreeView(v);
  ctx->views.clear();
}

namespace {

// Sparse-point seeding (patchmatch_stereo.cpp:499-533): seeds (u, v, depth)
// outside the image or with depth <= 0 are dropped, the rest fall into
// 32 x 32 pixel cells (column min(gx - 1, int(u) / 32)), and a cell's depth
// is the mean of its seeds, summed in seed order in FP64 as the reference
// does; stored as float, 0 for a cell without seeds.
void SeedCells(int w, int h, const double* seeds, int n, std::vector<float>* cells, int* gx, int* gy) {
  constexpr int kGrid = 32;
  *gx = (w + kGrid - 1) / kGrid;
  *gy = (h + kGrid - 1) / kGrid;
  const size_t nc = static_cast<size_t>(*gx) * *gy;
  std::vector<double> sum(nc, 0.0);
  std::vector<int> count(nc, 0);
  for (int i = 0; i < n; ++i) {
    const double su = seeds[3 * i], sv = seeds[3 * i + 1], sd = seeds[3 * i + 2];
    if (su < 0 || sv < 0 || su >= w || sv >= h) continue;
    if (sd <= 0.0) continue;
    const int cx = std::min(*gx - 1, static_cast<int>(su) / kGrid);
    const int cy = std::min(*gy - 1, static_cast<int>(sv) / kGrid);
    sum[static_cast<size_t>(cy) * *gx + cx] += sd;
    ++count[static_cast<size_t>(cy) * *gx + cx];
  }
  cells->assign(nc, 0.0f);
  for (size_t c = 0; c < nc; ++c)
    if (count[c] > 0) (*cells)[c] = static_cast<float>(sum[c] / static_cast<double>(count[c]));
}

int32_t RunPlanePatchmatch(recon_mvs_hip_context* ctx, int32_t ref_view, int32_t num_src,
                           const int32_t* src_views, const recon_mvs_plane_params* params,
                           const double* seed_points, int32_t num_seeds) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (CancelRequested(ctx)) return Cancelled(ctx);
  if (num_seeds < 0 || (num_seeds > 0 && !seed_points))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad seed points");
  PlaneArgs a;
  int32_t st = MakeArgs(ctx, ref_view, num_src, src_views, params, &a);
  if (st != RECON_MVS_OK) return st;
  st = EnsureState(ctx, ref_view);
  if (st != RECON_MVS_OK) return st;
  PlaneState& S = *ctx->states[ref_view];
  S.has_geo = false;
  a.plane = S.plane;
  a.prev = S.prev;
  a.cand = S.cand;
  a.cost = S.cost;
  a.support = S.support;
  if (num_seeds > 0) {
    // the staging vector of an earlier seeded call may still feed its copy
    if (!S.seed_host.empty()) RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int gx = 0, gy = 0;
    SeedCells(S.w, S.h, seed_points, num_seeds, &S.seed_host, &gx, &gy);
    const int nc = gx * gy;
    if (S.seed_capacity < nc) {
      if (S.seed_cells) (void)hipFree(S.seed_cells);
      S.seed_cells = nullptr;
      S.seed_capacity = 0;
      if (hipMalloc(&S.seed_cells, nc * sizeof(float)) != hipSuccess)
        return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "seed cell allocation failed");
      S.seed_capacity = nc;
    }
    RMVS_HIP(ctx, hipMemcpyAsync(S.seed_cells, S.seed_host.data(), nc * sizeof(float),
                                 hipMemcpyHostToDevice, ctx->stream));
    a.seed_cells = S.seed_cells;
    a.seed_gx = gx;
    a.seed_gy = gy;
  }
  int sched = 0;
  st = ResolveSchedule(ctx, a, params, &sched);
  if (st != RECON_MVS_OK) return st;
  std::vector<int> caps;
  st = ResolveCaps(ctx, a, params, sched, &caps);
  if (st != RECON_MVS_OK) return st;
  st = Launch(ctx, params, kInit, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched,
              caps[0]);
  if (st != RECON_MVS_OK) return st;
  // view-phased leading iterations (the 11x11 stride-2 window only);
  // automatic: none, unless the binned form is switched off, then the first
  // iteration for FP32 images (binned sweeps keep the unconverged gathers in
  // the caches better; FP16 / BF16 images are half the bytes, and their
  // fused sweeps already keep to the L2)
  int phased = 0;
  if (params->patch_size == 11 && params->window_stride == 2)
    phased = params->view_phased_iterations >= 0
                 ? params->view_phased_iterations
                 : (params->binned_phases == 0 && ctx->storage_format == RECON_MVS_PIXEL_F32 ? 1 : 0);
  for (int it = 0; it < params->iterations; ++it) {
    for (int color = 0; color < 2; ++color) {
      a.iter = it;
      a.color = color;
      if (it < phased)
        st = PhasedHalfSweep(ctx, params, a, ref_view, src_views, sched, it, color);
      else
        st = Launch(ctx, params, kSweep, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr, sched,
                    caps[1 + 2 * it + color]);
      if (st != RECON_MVS_OK) return st;
    }
  }
  return RECON_MVS_OK;
}

}  // namespace

extern "C" {

RECON_MVS_API void recon_mvs_plane_params_default(recon_mvs_plane_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->patch_size = 11;
  p->window_stride = 2;
  p->top_k = 4;
  p->iterations = 5;
  p->seed = 0;
  p->propagation_reach = 4;
  p->refine_iterations = 4;
  p->min_depth_m = 0.10;
  p->max_depth_m = 1.50;
  p->depth_perturb = 0.10;
  p->normal_perturb = 0.30;
  p->photo_threshold = 1.0;
  p->value_scale = 1.0;
  p->occupancy_cap = 0;
  p->capped_iterations = 1;
  p->view_phased_iterations = -1;
  p->binned_phases = -1;
}

RECON_MVS_API int32_t recon_mvs_hip_set_views(recon_mvs_hip_context* ctx, int32_t n,
                                              const recon_mvs_view* views, int32_t storage_format) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (n < 0 || (n > 0 && !views)) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad views");
  if (storage_format != RECON_MVS_PIXEL_F32 && storage_format != RECON_MVS_PIXEL_F16 &&
      storage_format != RECON_MVS_PIXEL_BF16 && storage_format != RECON_MVS_PIXEL_U8)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "storage format must be F32, F16, BF16 or U8");
  for (int i = 0; i < n; ++i) {
    const recon_mvs_view& v = views[i];
    if (!v.pixels || v.camera.width <= 0 || v.camera.height <= 0 ||
        (v.pixel_format != RECON_MVS_PIXEL_U8 && v.pixel_format != RECON_MVS_PIXEL_F32 &&
         v.pixel_format != RECON_MVS_PIXEL_BGRA8))
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad view");
    if (v.pixel_format == RECON_MVS_PIXEL_BGRA8 ? (v.row_stride != 0 && v.row_stride < 4 * v.camera.width)
                                                : v.row_stride != 0)
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad row_stride");
    // 8-bit storage holds 8-bit views (gray or B,G,R,A frames) exactly, and
    // nothing else
    if (storage_format == RECON_MVS_PIXEL_U8 && v.pixel_format == RECON_MVS_PIXEL_F32)
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "U8 storage needs 8-bit views");
    // sample addresses are formed in FP32: exact below 2^24 elements
    if (static_cast<int64_t>((v.camera.width + 1 + 63) & ~63) * v.camera.height >= (int64_t{1} << 24))
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT,
                  "view too large: (width + 1 rounded up to 64) x height must stay below 2^24");
  }
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  DestroyPlaneState(ctx);
  ctx->storage_format = storage_format;
  // any failure below leaves the context without views (never a view whose
  // pixels were not uploaded)
  auto fail = [&](int32_t code, const std::string& msg) {
    (void)hipStreamSynchronize(ctx->stream);
    DestroyPlaneState(ctx);
    return Fail(ctx, code, msg);
  };
  const size_t esz = storage_format == RECON_MVS_PIXEL_F32 ? 4 : storage_format == RECON_MVS_PIXEL_U8 ? 1 : 2;
  for (int i = 0; i < n; ++i) {
    const recon_mvs_view& in = views[i];
    auto* v = new ResidentView();
    v->w = in.camera.width;
    v->h = in.camera.height;
    v->pitch = (v->w + 1 + 63) & ~63;  // >= w + 1: the right tap of the last column is padding
    v->format = storage_format;
    v->cam = in.camera;
    ctx->views.push_back(v);
    const size_t bytes = static_cast<size_t>(v->pitch) * v->h * 2 * esz;
    if (hipMalloc(&v->pixels, bytes) != hipSuccess) {
      v->pixels = nullptr;
      return fail(RECON_MVS_ERR_OUT_OF_MEMORY, "view allocation failed");
    }
    const int src_stride = in.pixel_format == RECON_MVS_PIXEL_BGRA8
                               ? (in.row_stride ? in.row_stride : 4 * v->w)
                               : v->w * (in.pixel_format == RECON_MVS_PIXEL_U8 ? 1 : 4);
    const size_t in_bytes = static_cast<size_t>(src_stride) * v->h;
    void* staging = Ensure(ctx, kSlotQueryImgA, in_bytes);
    if (!staging) return fail(RECON_MVS_ERR_OUT_OF_MEMORY, "staging allocation failed");
    hipError_t e = hipMemcpyAsync(staging, in.pixels, in_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return fail(RECON_MVS_ERR_GPU_FAILURE, std::string("view upload: ") + hipGetErrorString(e));
    const int total = v->pitch * v->h;
    const dim3 g((total + 255) / 256), b(256);
    if (storage_format == RECON_MVS_PIXEL_F32)
      hipLaunchKernelGGL(k_store_view<RECON_MVS_PIXEL_F32>, g, b, 0, ctx->stream, staging, in.pixel_format, src_stride, v->w, v->h, v->pitch, v->pixels);
    else if (storage_format == RECON_MVS_PIXEL_F16)
      hipLaunchKernelGGL(k_store_view<RECON_MVS_PIXEL_F16>, g, b, 0, ctx->stream, staging, in.pixel_format, src_stride, v->w, v->h, v->pitch, v->pixels);
    else if (storage_format == RECON_MVS_PIXEL_BF16)
      hipLaunchKernelGGL(k_store_view<RECON_MVS_PIXEL_BF16>, g, b, 0, ctx->stream, staging, in.pixel_format, src_stride, v->w, v->h, v->pitch, v->pixels);
    else
      hipLaunchKernelGGL(k_store_view<RECON_MVS_PIXEL_U8>, g, b, 0, ctx->stream, staging, in.pixel_format, src_stride, v->w, v->h, v->pitch, v->pixels);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(RECON_MVS_ERR_GPU_FAILURE, std::string("view store: ") + hipGetErrorString(e));
    // staging is reused by the next view: finish this one first
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return fail(RECON_MVS_ERR_GPU_FAILURE, std::string("view upload: ") + hipGetErrorString(e));
  }
  ctx->states.assign(ctx->views.size(), nullptr);
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_bgra_to_gray(recon_mvs_hip_context* ctx, const uint8_t* bgra,
                                                 int32_t width, int32_t height, int32_t stride,
                                                 uint8_t* gray) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!bgra || !gray || width <= 0 || height <= 0 || stride < 4 * width)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bgra_to_gray: bad arguments");
  const size_t in_bytes = static_cast<size_t>(stride) * height, out_bytes = static_cast<size_t>(width) * height;
  uint8_t* d_in = static_cast<uint8_t*>(Ensure(ctx, kSlotQueryImgA, in_bytes));
  uint8_t* d_out = static_cast<uint8_t*>(Ensure(ctx, kSlotQueryImgB, out_bytes));
  if (!d_in || !d_out) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  RMVS_HIP(ctx, hipMemcpyAsync(d_in, bgra, in_bytes, hipMemcpyHostToDevice, ctx->stream));
  const int n = width * height;
  hipLaunchKernelGGL(k_bgra_to_gray, dim3((n + 255) / 256), dim3(256), 0, ctx->stream, d_in, width, height,
                     stride, d_out);
  RMVS_LAUNCH_CHECK(ctx);
  RMVS_HIP(ctx, hipMemcpyAsync(gray, d_out, out_bytes, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_plane_patchmatch(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                     int32_t num_src, const int32_t* src_views,
                                                     const recon_mvs_plane_params* params) {
  return RunPlanePatchmatch(ctx, ref_view, num_src, src_views, params, nullptr, 0);
}

RECON_MVS_API int32_t recon_mvs_hip_plane_patchmatch_seeded(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                            int32_t num_src, const int32_t* src_views,
                                                            const recon_mvs_plane_params* params,
                                                            const double* seed_points, int32_t num_seeds) {
  return RunPlanePatchmatch(ctx, ref_view, num_src, src_views, params, seed_points, num_seeds);
}

RECON_MVS_API int32_t recon_mvs_hip_plane_refine(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                 int32_t num_src, const int32_t* src_views,
                                                 const recon_mvs_plane_params* params) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (CancelRequested(ctx)) return Cancelled(ctx);
  PlaneArgs a;
  int32_t st = MakeArgs(ctx, ref_view, num_src, src_views, params, &a);
  if (st != RECON_MVS_OK) return st;
  if (ref_view >= static_cast<int>(ctx->states.size()) || !ctx->states[ref_view])
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "no depth map in this slot");
  PlaneState& S = *ctx->states[ref_view];
  a.plane = S.plane;
  a.cost = S.cost;
  a.support = S.support;
  const double baseline = MeanBaselineViews(ctx, ref_view, num_src, src_views);
  const ResidentView& R = *ctx->views[ref_view];
  const double focal = 0.5 * (R.cam.fx + R.cam.fy);
  if (!(baseline > 1e-9)) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "zero baseline");
  const float inv_per_px = static_cast<float>(1.0 / (focal * baseline));
  return Launch(ctx, params, kRefine, a, inv_per_px, params->refine_iterations, 0, nullptr, nullptr,
                nullptr, nullptr, nullptr, std::max(0, params->tile_schedule));
}

RECON_MVS_API int32_t recon_mvs_hip_set_plane_state(recon_mvs_hip_context* ctx, int32_t view,
                                                    int32_t num_src, const int32_t* src_views,
                                                    const recon_mvs_plane_params* params,
                                                    const float* depth, const float* normals) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!depth || !normals) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null state");
  PlaneArgs a;
  int32_t st = MakeArgs(ctx, view, num_src, src_views, params, &a);
  if (st != RECON_MVS_OK) return st;
  st = EnsureState(ctx, view);
  if (st != RECON_MVS_OK) return st;
  PlaneState& S = *ctx->states[view];
  const size_t P = static_cast<size_t>(S.w) * S.h;
  std::vector<float4> host(P);
  for (size_t i = 0; i < P; ++i)
    host[i] = make_float4(depth[i], normals[3 * i], normals[3 * i + 1], normals[3 * i + 2]);
  RMVS_HIP(ctx, hipMemcpyAsync(S.plane, host.data(), P * sizeof(float4), hipMemcpyHostToDevice, ctx->stream));
  a.plane = S.plane;
  a.cost = S.cost;
  a.support = S.support;
  st = Launch(ctx, params, kScore, a, 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
              std::max(0, params->tile_schedule));
  if (st != RECON_MVS_OK) return st;
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  S.has_geo = false;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_geo_consistency(recon_mvs_hip_context* ctx, int32_t num_maps,
                                                    const int32_t* views, double threshold,
                                                    int32_t min_views) {
  (void)min_views;
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (CancelRequested(ctx)) return Cancelled(ctx);
  if (num_maps < 1 || num_maps > kMaxGeo || !views)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "num_maps out of [1, 16]");
  GeoArgs g;
  std::memset(&g, 0, sizeof(g));
  g.nmaps = num_maps;
  g.thr = threshold;
  for (int m = 0; m < num_maps; ++m) {
    const int v = views[m];
    if (v < 0 || v >= static_cast<int>(ctx->states.size()) || !ctx->states[v])
      return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "view without a depth map");
    const PlaneState& S = *ctx->states[v];
    g.v[m].w = S.w;
    g.v[m].h = S.h;
    g.v[m].plane = S.plane;
    g.v[m].cost = S.cost;
    g.v[m].cam = ctx->views[v]->cam;
  }
  for (int m = 0; m < num_maps; ++m) {
    PlaneState& S = *ctx->states[views[m]];
    g.self = m;
    g.geo_support = S.geo_support;
    hipEvent_t eb, ee;
    TimerBegin(ctx, kGeo, &eb, &ee);
    hipLaunchKernelGGL(k_geo_consistency, dim3((S.w * S.h + 255) / 256), dim3(256), 0, ctx->stream, g);
    TimerEnd(ctx, kGeo, eb, ee);
    RMVS_LAUNCH_CHECK(ctx);
    S.has_geo = true;
  }
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_get_depth_map(recon_mvs_hip_context* ctx, int32_t view,
                                                  int32_t drop_inconsistent, float* depth,
                                                  float* normals, float* cost, uint8_t* support,
                                                  float* confidence, uint8_t* geo_support) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (view < 0 || view >= static_cast<int>(ctx->states.size()) || !ctx->states[view])
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "no depth map in this slot");
  PlaneState& S = *ctx->states[view];
  const size_t P = static_cast<size_t>(S.w) * S.h;
  float* d_depth = static_cast<float*>(Ensure(ctx, kSlotOutDepth, P * 4));
  float* d_conf = static_cast<float*>(Ensure(ctx, kSlotOutConf, P * 4));
  if (!d_depth || !d_conf) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  const int min_views = drop_inconsistent ? drop_inconsistent : 0;
  hipLaunchKernelGGL(k_plane_output, dim3((P + 255) / 256), dim3(256), 0, ctx->stream, static_cast<int>(P),
                     S.plane, S.cost, (drop_inconsistent && S.has_geo) ? S.geo_support : nullptr,
                     min_views, d_depth, d_conf);
  RMVS_LAUNCH_CHECK(ctx);
  if (depth) RMVS_HIP(ctx, hipMemcpyAsync(depth, d_depth, P * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (confidence) RMVS_HIP(ctx, hipMemcpyAsync(confidence, d_conf, P * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (cost) RMVS_HIP(ctx, hipMemcpyAsync(cost, S.cost, P * 4, hipMemcpyDeviceToHost, ctx->stream));
  if (support) RMVS_HIP(ctx, hipMemcpyAsync(support, S.support, P, hipMemcpyDeviceToHost, ctx->stream));
  if (geo_support) {
    if (!S.has_geo) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "consistency filter has not run");
    RMVS_HIP(ctx, hipMemcpyAsync(geo_support, S.geo_support, P, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (normals) {
    std::vector<float4> host(P);
    RMVS_HIP(ctx, hipMemcpyAsync(host.data(), S.plane, P * sizeof(float4), hipMemcpyDeviceToHost, ctx->stream));
    RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < P; ++i) {
      normals[3 * i] = host[i].y;
      normals[3 * i + 1] = host[i].z;
      normals[3 * i + 2] = host[i].w;
    }
  }
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_eval_plane_costs(recon_mvs_hip_context* ctx, int32_t ref_view,
                                                     int32_t num_src, const int32_t* src_views,
                                                     const recon_mvs_plane_params* params, int32_t n,
                                                     const float* hyps, float* cost,
                                                     int32_t* valid_views, float* view_ncc) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (n < 0 || (n > 0 && (!hyps || !cost || !valid_views)))
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "bad query arrays");
  PlaneArgs a;
  int32_t st = MakeArgs(ctx, ref_view, num_src, src_views, params, &a);
  if (st != RECON_MVS_OK) return st;
  if (n == 0) return RECON_MVS_OK;
  float* dh = static_cast<float*>(Ensure(ctx, kSlotQueryIn, sizeof(float) * 6 * n));
  float* dc = static_cast<float*>(Ensure(ctx, kSlotQueryOut, sizeof(float) * n * (2 + num_src)));
  if (!dh || !dc) return Fail(ctx, RECON_MVS_ERR_OUT_OF_MEMORY, "alloc failed");
  int* dv = reinterpret_cast<int*>(dc + n);
  float* dn = dc + 2 * static_cast<size_t>(n);
  RMVS_HIP(ctx, hipMemcpyAsync(dh, hyps, sizeof(float) * 6 * n, hipMemcpyHostToDevice, ctx->stream));
  st = Launch(ctx, params, kEval, a, 0.0f, 0, n, dh, dc, dv, view_ncc ? dn : nullptr, nullptr,
              std::max(0, params->tile_schedule));
  if (st != RECON_MVS_OK) return st;
  RMVS_HIP(ctx, hipMemcpyAsync(cost, dc, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipMemcpyAsync(valid_views, dv, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
  if (view_ncc)
    RMVS_HIP(ctx, hipMemcpyAsync(view_ncc, dn, sizeof(float) * n * num_src, hipMemcpyDeviceToHost, ctx->stream));
  RMVS_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_timing(recon_mvs_hip_context* ctx, int32_t enable, int32_t reset,
                                           double* ms, int64_t* counts) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  TimerCollect(ctx);
  if (ms)
    for (int i = 0; i < 8; ++i) ms[i] = ctx->timed_ms[i];
  if (counts)
    for (int i = 0; i < 8; ++i) counts[i] = ctx->timed_count[i];
  if (reset)
    for (int i = 0; i < 8; ++i) {
      ctx->timed_ms[i] = 0.0;
      ctx->timed_count[i] = 0;
    }
  ctx->timing = enable != 0;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_autotune(recon_mvs_hip_context* ctx, int32_t ref_view,
                                             int32_t num_src, const int32_t* src_views,
                                             const recon_mvs_plane_params* params,
                                             int32_t* best_schedule, double* ms_per_schedule) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!best_schedule) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null best_schedule");
  PlaneArgs a;
  int32_t st = MakeArgs(ctx, ref_view, num_src, src_views, params, &a);
  if (st != RECON_MVS_OK) return st;
  st = EnsureState(ctx, ref_view);
  if (st != RECON_MVS_OK) return st;
  PlaneState& S = *ctx->states[ref_view];
  a.plane = S.plane;
  a.prev = S.prev;
  a.cand = S.cand;
  a.cost = S.cost;
  a.support = S.support;
  int best = 0;
  st = Autotune(ctx, a, params, &best, ms_per_schedule);
  if (st != RECON_MVS_OK) return st;
  ctx->tuned_schedule[TuneKey(ctx, a, params)] = best;
  *best_schedule = best;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_batch_create(recon_mvs_hip_context* ctx, int32_t num_jobs,
                                                 const recon_mvs_batch_job* jobs,
                                                 const recon_mvs_plane_params* params,
                                                 recon_mvs_hip_batch** out) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!out || !jobs || !params || num_jobs < 1)
    return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "batch_create: bad arguments");
  *out = nullptr;
  std::vector<PlaneArgs> args(num_jobs);
  std::vector<int> seen(ctx->views.size(), 0);
  for (int j = 0; j < num_jobs; ++j) {
    int32_t st = MakeArgs(ctx, jobs[j].ref_view, jobs[j].num_src, jobs[j].src_views, params, &args[j]);
    if (st != RECON_MVS_OK) return st;
    if (seen[jobs[j].ref_view]++) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "a view is the reference of two jobs");
    st = EnsureState(ctx, jobs[j].ref_view);
    if (st != RECON_MVS_OK) return st;
    PlaneState& S = *ctx->states[jobs[j].ref_view];
    S.has_geo = false;
    args[j].plane = S.plane;
    args[j].prev = S.prev;
    args[j].cand = S.cand;
    args[j].cost = S.cost;
    args[j].support = S.support;
  }
  // work units, most views first (largest units leave the queues first)
  std::vector<int> order(num_jobs);
  for (int j = 0; j < num_jobs; ++j) order[j] = j;
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return args[x].nsrc > args[y].nsrc; });
  const int nph = 1 + 2 * params->iterations;
  // candidate tile schedules: the given one, or every schedule (those not
  // compiled for this window / format fail to launch and drop out)
  std::vector<int> scheds;
  if (params->tile_schedule >= 0) scheds.push_back(params->tile_schedule);
  else for (int sc = 0; sc < kNumSchedules; ++sc) scheds.push_back(sc);
  size_t max_units = 0;
  for (int sc : scheds) {
    std::vector<int2> u;
    std::vector<int> q;
    BatchUnits(&args, order, sc, &u, &q);
    max_units = std::max(max_units, u.size());
  }
  auto* b = new recon_mvs_hip_batch();
  b->params = *params;
  for (int j = 0; j < num_jobs; ++j) b->refs.push_back(jobs[j].ref_view);
  auto cleanup = [&](int32_t code, const char* msg) {
    if (b->d_args) (void)hipFree(b->d_args);
    if (b->d_units) (void)hipFree(b->d_units);
    if (b->d_qoff) (void)hipFree(b->d_qoff);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
    return Fail(ctx, code, msg);
  };
  if (hipMalloc(&b->d_args, sizeof(PlaneArgs) * num_jobs) != hipSuccess ||
      hipMalloc(&b->d_units, sizeof(int2) * max_units) != hipSuccess ||
      hipMalloc(&b->d_qoff, sizeof(int) * (kXcds + 1)) != hipSuccess ||
      hipMalloc(&b->d_counter, sizeof(unsigned int) * kXcds * kCounterStride) != hipSuccess)
    return cleanup(RECON_MVS_ERR_OUT_OF_MEMORY, "batch allocation failed");
  std::vector<int2> units;
  std::vector<int> qoff;
  auto upload = [&](int sc) -> bool {
    BatchUnits(&args, order, sc, &units, &qoff);
    return hipMemcpy(b->d_args, args.data(), sizeof(PlaneArgs) * num_jobs, hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(b->d_units, units.data(), sizeof(int2) * units.size(), hipMemcpyHostToDevice) == hipSuccess &&
           hipMemcpy(b->d_qoff, qoff.data(), sizeof(int) * qoff.size(), hipMemcpyHostToDevice) == hipSuccess;
  };
  // configuration per phase (index 0: init): occupancy cap none
  // (occupancy_cap < 0), the given cap in the leading capped phases (> 0),
  // or tuned (0): every candidate schedule's phases are timed on the batch
  // itself uncapped and at 2 workgroups per CU, fused and (binned_phases =
  // -1) binned, each phase keeps its fastest, and the schedule with the
  // least total wins; binned_phases > 0 forces the binned form on the
  // leading phases.  Results do not depend on any of these choices.
  int sched = scheds[0];
  std::vector<int> caps(nph, 0);
  if (params->occupancy_cap > 0) {
    const int ci = params->capped_iterations < 0 ? 1 : params->capped_iterations;
    for (int ph = 0; ph < std::min(nph, 1 + 2 * ci); ++ph) caps[ph] = params->occupancy_cap;
  }
  const int forced = HasBinned(params) && params->binned_phases > 0 ? std::min(nph, params->binned_phases) : 0;
  for (int ph = 0; ph < forced; ++ph) caps[ph] = kBinnedCfg;
  const bool tune = params->occupancy_cap == 0 || scheds.size() > 1;
  if (tune) {
    const bool timing = ctx->timing;
    ctx->timing = false;
    double best_total = 1e300;
    std::vector<int> trial_cfgs = {0, 2};
    if (HasBinned(params) && params->binned_phases < 0) {
      trial_cfgs.push_back(kBinnedCfg);
      trial_cfgs.push_back(kBinnedCfg | 2);
    }
    for (int sc : scheds) {
      if (!upload(sc)) {
        ctx->timing = timing;
        return cleanup(RECON_MVS_ERR_GPU_FAILURE, "batch upload failed");
      }
      BatchLaunch bl{b->d_args, b->d_units, b->d_qoff, b->d_counter, 0, ctx->cancel_dev};
      std::vector<float> ms, best(nph, 3.0e38f);
      std::vector<int> cap_of(nph, 0);
      // a first run loads the code; a schedule not compiled here drops out
      if (TimeBatchPhases(ctx, params, args[0], bl, sc, std::vector<int>(nph, 0), &ms) != RECON_MVS_OK) {
        ctx->error.clear();
        continue;
      }
      int32_t st = RECON_MVS_OK;
      const size_t ncfg = params->occupancy_cap == 0 ? trial_cfgs.size() : 1;
      for (size_t t = 0; t < ncfg && st == RECON_MVS_OK; ++t) {
        std::vector<int> trial = params->occupancy_cap == 0 ? std::vector<int>(nph, trial_cfgs[t]) : caps;
        for (int ph = 0; ph < forced; ++ph) trial[ph] = kBinnedCfg;
        st = TimeBatchPhases(ctx, params, args[0], bl, sc, trial, &ms);
        for (int ph = 0; ph < nph && st == RECON_MVS_OK; ++ph)
          if (ms[ph] < best[ph]) {
            best[ph] = ms[ph];
            cap_of[ph] = trial[ph];
          }
      }
      if (st != RECON_MVS_OK) {
        ctx->timing = timing;
        return cleanup(st, "batch tuning failed");
      }
      double total = 0.0;
      for (float x : best) total += x;
      if (total < best_total) {
        best_total = total;
        sched = sc;
        caps = cap_of;
      }
    }
    ctx->timing = timing;
    if (best_total >= 1e300)
      return cleanup(RECON_MVS_ERR_NOT_IMPLEMENTED, "no tile schedule compiled for this window / format");
  }
  if (!upload(sched)) return cleanup(RECON_MVS_ERR_GPU_FAILURE, "batch upload failed");
  b->schedule = sched;
  b->num_units = static_cast<int>(units.size());
  // capture init + all half-sweeps as one graph
  const bool timing = ctx->timing;
  ctx->timing = false;
  if (hipStreamSynchronize(ctx->stream) != hipSuccess ||
      hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    ctx->timing = timing;
    return cleanup(RECON_MVS_ERR_GPU_FAILURE, "graph capture failed to start");
  }
  int32_t st = RECON_MVS_OK;
  for (int phase = -1; phase < 2 * params->iterations && st == RECON_MVS_OK; ++phase) {
    BatchLaunch bl{b->d_args, b->d_units, b->d_qoff, b->d_counter, phase, ctx->cancel_dev};
    st = Launch(ctx, params, kBatch, args[0], 0.0f, 0, 0, nullptr, nullptr, nullptr, nullptr, &bl, sched,
                caps[phase + 1]);
  }
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &graph);
  ctx->timing = timing;
  if (st != RECON_MVS_OK || ec != hipSuccess || !graph) {
    if (graph) (void)hipGraphDestroy(graph);
    return cleanup(RECON_MVS_ERR_GPU_FAILURE, "graph capture failed");
  }
  const hipError_t ei = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (ei != hipSuccess) return cleanup(RECON_MVS_ERR_GPU_FAILURE, "graph instantiation failed");
  *out = b;
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_batch_run(recon_mvs_hip_context* ctx, recon_mvs_hip_batch* b) {
  if (!ctx) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (!b) return Fail(ctx, RECON_MVS_ERR_INVALID_ARGUMENT, "null batch");
  if (CancelRequested(ctx)) return Cancelled(ctx);
  hipEvent_t eb, ee;
  TimerBegin(ctx, kBatch, &eb, &ee);
  RMVS_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
  TimerEnd(ctx, kBatch, eb, ee);
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_hip_batch_schedule(const recon_mvs_hip_batch* b) {
  return b ? b->schedule : -1;
}

RECON_MVS_API int32_t recon_mvs_hip_batch_destroy(recon_mvs_hip_context* ctx, recon_mvs_hip_batch* b) {
  if (!b) return RECON_MVS_OK;
  if (ctx) (void)hipStreamSynchronize(ctx->stream);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->d_args) (void)hipFree(b->d_args);
  if (b->d_units) (void)hipFree(b->d_units);
  if (b->d_qoff) (void)hipFree(b->d_qoff);
  if (b->d_counter) (void)hipFree(b->d_counter);
  delete b;
  return RECON_MVS_OK;
}

}  // extern "C"
