// test_mvs_api.cpp — a caller of the recon:: MVS API (include/recon/mvs.h),
// built against librecon_mvs_hip.so exactly as a caller of the reference's
// mvs.h would be (tests/unit/test_mvs.cpp is the reference's such caller).
// It runs the reference's unit-test scene through PatchMatchStereo ->
// RefineSubpixel -> FuseDepthMaps on the GPU and applies the thresholds of
// tests/unit/test_mvs.cpp:344-465.  Exit code 0 = all checks passed.
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "recon/camera.h"
#include "recon/mvs.h"
#include "recon/mvs_cache.h"
#include "recon/mvs_gpu.h"
#include "recon_mvs_hip.h"

extern "C" int32_t recon_scene_reference_plane(int32_t num_views, double baseline_m,
                                               recon_mvs_camera* cams, uint8_t* images,
                                               float* truth_depth);

static int g_fail = 0;
static void Expect(bool ok, const char* what) {
  std::printf("  [%s] %s\n", ok ? "ok" : "FAIL", what);
  if (!ok) ++g_fail;
}

int main() {
  const int nv = 5, W = 160, H = 120;
  std::vector<recon_mvs_camera> cams(nv);
  std::vector<uint8_t> imgs(static_cast<size_t>(nv) * W * H);
  std::vector<float> truth(static_cast<size_t>(W) * H);
  if (recon_scene_reference_plane(nv, 0.016, cams.data(), imgs.data(), truth.data()) != 0) return 2;

  std::vector<recon::CameraPose> poses(nv);
  recon::CameraIntrinsics intr;
  intr.fx = cams[0].fx; intr.fy = cams[0].fy; intr.cx = cams[0].cx; intr.cy = cams[0].cy;
  intr.width = W; intr.height = H;
  for (int v = 0; v < nv; ++v) {
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) poses[v].R(i, j) = cams[v].R[3 * i + j];
      poses[v].t(i) = cams[v].t[i];
    }
  }

  // camera model round trip with distortion (test_mvs.cpp:210-253 checks the same property)
  {
    recon::CameraIntrinsics k;
    k.fx = k.fy = 1250.0; k.cx = 960.0; k.cy = 540.0; k.width = 1920; k.height = 1080;
    k.k1 = -0.22; k.k2 = 0.06; k.k3 = -0.005; k.p1 = 0.0008; k.p2 = -0.0006;
    double worst = 0.0;
    for (int i = 0; i < 150; ++i) {
      const double u = 40.0 + 9.0 * i, v = 30.0 + 5.0 * i;
      recon::Vec3d r;
      k.PixelToRay(u, v, &r);
      double u2, v2;
      k.ProjectPoint(r, &u2, &v2);
      worst = std::max(worst, std::hypot(u2 - u, v2 - v));
    }
    Expect(worst < 1e-6, "distortion round trip < 1e-6 px");
    double a, b;
    Expect(!k.ProjectPoint(recon::Vec3d(0.0, 0.0, -1.0), &a, &b), "point behind camera rejected");
  }

  // NCC of a window with itself is 1 (GPU batch entry point)
  {
    const uint8_t* a = imgs.data();
    const recon::NccResult r = recon::ComputeNcc(a, W, H, a, W, H, 80.0, 60.0, 80.0, 60.0, 11);
    Expect(r.valid && std::fabs(r.ncc - 1.0) < 1e-6 && !r.weak_texture, "ComputeNcc(a, a) == 1");
  }

  // view selection
  {
    recon::MvsConfig c;
    c.num_views = 3;
    const std::vector<int> sel = recon::SelectSourceViews(poses[0], poses, intr, c);
    bool ok = sel.size() <= 3;
    for (size_t i = 0; i < sel.size(); ++i) {
      ok = ok && sel[i] != 0;
      for (size_t j = i + 1; j < sel.size(); ++j) ok = ok && sel[i] != sel[j];
    }
    Expect(ok, "SelectSourceViews: <= N, no self, no duplicates");
  }

  recon::MvsConfig cfg;
  cfg.num_views = nv - 1;
  cfg.num_threads = 8;
  cfg.patch_size = 11;
  cfg.iterations = 3;
  cfg.random_search_count = 5;
  cfg.random_search_radius_ratio = 0.35;
  cfg.subpixel_iterations = 5;
  cfg.min_depth_m = 0.15;
  cfg.max_depth_m = 0.80;
  cfg.fusion_voxel_size_m = 0.0005;
  std::vector<int> sources = {1, 2, 3, 4};
  std::vector<const uint8_t*> src;
  std::vector<int> sw, sh;
  std::vector<recon::CameraIntrinsics> src_intr;
  for (int s : sources) {
    src.push_back(imgs.data() + static_cast<size_t>(s) * W * H);
    sw.push_back(W);
    sh.push_back(H);
    src_intr.push_back(intr);
  }
  recon::DepthMap depth;
  const uint32_t filled = recon::PatchMatchStereo(imgs.data(), W, H, src, sw, sh, poses[0], poses,
                                                  intr, src_intr, sources, cfg, nullptr, &depth);
  std::printf("  filled %u / %d\n", filled, W * H);
  Expect(filled > W * H / 2, "PatchMatchStereo fills more than half");
  int multi = 0;
  for (size_t i = 0; i < depth.depth.size(); ++i)
    if (std::isfinite(depth.depth[i]) && depth.support_count[i] >= 2) ++multi;
  Expect(multi > static_cast<int>(filled / 2), "support >= 2 on most pixels");
  const uint32_t refined = recon::RefineSubpixel(imgs.data(), W, H, src, sw, sh, poses[0], poses,
                                                 intr, src_intr, sources, cfg, &depth);
  std::printf("  refined %u\n", refined);
  std::vector<double> err;
  for (int y = 20; y < H - 20; ++y)
    for (int x = 20; x < W - 20; ++x) {
      const float d = depth.depth[static_cast<size_t>(y) * W + x];
      const float t = truth[static_cast<size_t>(y) * W + x];
      if (std::isfinite(d) && t > 0.0f) err.push_back(std::fabs(double(d) - t) * 1000.0);
    }
  std::sort(err.begin(), err.end());
  const double median = err.empty() ? 1e9 : err[err.size() / 2];
  const double p95 = err.empty() ? 1e9 : err[static_cast<size_t>(err.size() * 0.95)];
  std::printf("  median %.4f mm  p95 %.4f mm  n=%zu\n", median, p95, err.size());
  Expect(median <= 19.1 * 0.15 * 1.5, "median depth error within 1.5x theory");
  Expect(p95 <= 19.1 * 0.15 * 3.0, "p95 depth error within 3x theory");
  const recon::PointCloud cloud = recon::FuseDepthMaps({&depth}, {poses[0]}, {intr}, cfg);
  std::printf("  fused %zu points\n", cloud.size());
  Expect(cloud.size() > 1000, "FuseDepthMaps produces > 1000 points");
  // the cloud in Poisson input form: weights are the confidences clamped at 0
  {
    const recon::OrientedPointCloud oc = recon::ToOrientedCloud(cloud);
    bool ok = oc.size() == cloud.size() && oc.weights.size() == cloud.size() &&
              oc.colors.size() == cloud.colors.size();
    for (size_t i = 0; ok && i < cloud.size(); ++i)
      ok = oc.weights[i] == std::max(0.0f, cloud.confidences[i]) &&
           oc.positions[i](0) == cloud.positions[i](0) && oc.normals[i](2) == cloud.normals[i](2);
    Expect(ok, "ToOrientedCloud: weights = max(0, confidence)");
  }
  // depth-map spill: written, read back bit-exact, a corrupted file rejected and deleted
  {
    char tmpl[] = "/tmp/recon_spill_XXXXXX";
    const char* dir = mkdtemp(tmpl);
    const std::string sha = dir ? recon::SpillDepthMap(dir, 7, depth) : std::string();
    Expect(sha.size() == 64, "SpillDepthMap returns a sha256");
    recon::DepthMap back;
    const bool loaded = dir && recon::LoadDepthMap(dir, 7, &back);
    Expect(loaded && back.width == W && back.height == H &&
               std::memcmp(back.depth.data(), depth.depth.data(), 4 * depth.depth.size()) == 0 &&
               back.support_count == depth.support_count && back.confidence == depth.confidence,
           "LoadDepthMap round trip is bit-exact");
    if (dir) {
      const std::string f = std::string(dir) + "/mvs_depth/ref_00007.f32";
      FILE* fp = std::fopen(f.c_str(), "r+b");
      if (fp) {
        std::fseek(fp, 100, SEEK_SET);
        std::fputc(0x5a, fp);
        std::fclose(fp);
      }
      recon::DepthMap bad;
      Expect(!recon::LoadDepthMap(dir, 7, &bad) && std::fopen(f.c_str(), "rb") == nullptr,
             "corrupted spill rejected and deleted");
      std::remove((std::string(dir) + "/cache_manifest.json").c_str());
      rmdir((std::string(dir) + "/mvs_depth").c_str());
      rmdir(dir);
    }
  }
  // index contract: ambiguous intrinsics length returns 0
  {
    recon::DepthMap d2;
    std::vector<recon::CameraIntrinsics> bad(3, intr);
    Expect(recon::PatchMatchStereo(imgs.data(), W, H, src, sw, sh, poses[0], poses, intr, bad,
                                   sources, cfg, nullptr, &d2) == 0,
           "ambiguous intrinsics -> 0");
  }
  // GPU controls (recon/mvs_gpu.h): a cancelled PatchMatchStereo produces
  // nothing (the reference's skip value 0) and reports status 5; a callback
  // cancels at its own poll; clearing restores normal runs
  {
    recon::ClearMvsGpuStatus();
    recon::CancelMvs();
    Expect(recon::MvsCancelled(), "CancelMvs sets the cancelled state");
    recon::DepthMap dc;
    Expect(recon::PatchMatchStereo(imgs.data(), W, H, src, sw, sh, poses[0], poses, intr, src_intr,
                                   sources, cfg, nullptr, &dc) == 0 &&
               recon::LastMvsGpuStatus() == 5,
           "cancelled PatchMatchStereo -> 0, LastMvsGpuStatus 5");
    recon::ClearMvsCancel();
    Expect(!recon::MvsCancelled(), "ClearMvsCancel clears it");
    int polls = 0;
    recon::SetMvsCancelCallback([](void* u) { return ++*static_cast<int*>(u) >= 3 ? 1 : 0; }, &polls);
    Expect(recon::PatchMatchStereo(imgs.data(), W, H, src, sw, sh, poses[0], poses, intr, src_intr,
                                   sources, cfg, nullptr, &dc) == 0 && polls == 3,
           "the callback cancels at its third poll (the third directional pass)");
    recon::SetMvsCancelCallback(nullptr, nullptr);
    recon::ClearMvsCancel();
    recon::ClearMvsGpuStatus();
    Expect(recon::PatchMatchStereo(imgs.data(), W, H, src, sw, sh, poses[0], poses, intr, src_intr,
                                   sources, cfg, nullptr, &dc) > 0 &&
               recon::LastMvsGpuStatus() == 0,
           "after clearing, PatchMatchStereo runs");
  }
  std::printf(g_fail ? "FAILED %d\n" : "ALL PASSED\n", g_fail);
  return g_fail ? 1 : 0;
}
