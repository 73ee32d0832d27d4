// spill.cpp — depth maps spilled to and restored from the session cache, as
// the reference's memory design specifies them (docs/13_内存与降级.md §3.3.2-
// §3.3.3): `<cache>/mvs_depth/ref_%05d.f32` holding the depth with its
// confidence and support counts, written to `*.tmp` then renamed (atomic),
// and recorded in `<cache>/cache_manifest.json` with its byte count and
// SHA-256; a read verifies the SHA-256 and reports a mismatch so the caller
// recomputes the map.
//
// File layout (little endian): a 32-byte header {char magic[4] = "RMVD";
// uint32 version = 1; int32 width, height, ref_index; uint32 planes = 3;
// uint32 reserved[2]} followed by three float32 planes of width x height:
// depth (NaN = invalid), confidence, support count.
#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "recon_mvs_hip.h"

namespace {

// ---- SHA-256 (FIPS 180-4) ----
struct Sha256 {
  uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  uint8_t buf[64];
  size_t fill = 0;
  uint64_t bits = 0;

  static uint32_t Rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

  void Block(const uint8_t* p) {
    static const uint32_t k[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u,
        0xab1c5ed5u, 0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu,
        0x9bdc06a7u, 0xc19bf174u, 0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu,
        0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau, 0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u,
        0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u, 0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu,
        0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u, 0xa2bfe8a1u, 0xa81a664bu,
        0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u, 0x19a4c116u,
        0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
        0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u,
        0xc67178f2u};
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
      w[i] = (uint32_t{p[4 * i]} << 24) | (uint32_t{p[4 * i + 1]} << 16) | (uint32_t{p[4 * i + 2]} << 8) |
             uint32_t{p[4 * i + 3]};
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = Rotr(w[i - 15], 7) ^ Rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      const uint32_t s1 = Rotr(w[i - 2], 17) ^ Rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t S1 = Rotr(e, 6) ^ Rotr(e, 11) ^ Rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + k[i] + w[i];
      const uint32_t S0 = Rotr(a, 2) ^ Rotr(a, 13) ^ Rotr(a, 22);
      const uint32_t mj = (a & b) ^ (a & c) ^ (b & c);
      const uint32_t t2 = S0 + mj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }

  void Update(const void* data, size_t n) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    bits += static_cast<uint64_t>(n) * 8;
    while (n > 0) {
      const size_t take = std::min(n, 64 - fill);
      std::memcpy(buf + fill, p, take);
      fill += take;
      p += take;
      n -= take;
      if (fill == 64) {
        Block(buf);
        fill = 0;
      }
    }
  }

  std::string HexDigest() {
    const uint64_t len = bits;
    const uint8_t one = 0x80, zero = 0;
    Update(&one, 1);
    while (fill != 56) Update(&zero, 1);
    uint8_t lb[8];
    for (int i = 0; i < 8; ++i) lb[i] = static_cast<uint8_t>(len >> (56 - 8 * i));
    Update(lb, 8);
    char out[65];
    for (int i = 0; i < 8; ++i) std::snprintf(out + 8 * i, 9, "%08x", h[i]);
    return std::string(out, 64);
  }
};

struct Header {
  char magic[4];
  uint32_t version;
  int32_t width, height, ref_index;
  uint32_t planes;
  uint32_t reserved[2];
};
static_assert(sizeof(Header) == 32, "spill header is 32 bytes");

std::string MapName(int32_t ref_index) {
  char name[32];
  std::snprintf(name, sizeof(name), "mvs_depth/ref_%05d.f32", ref_index);
  return name;
}

bool MakeDir(const std::string& d) { return mkdir(d.c_str(), 0755) == 0 || errno == EEXIST; }

bool ReadFile(const std::string& path, std::vector<uint8_t>* out) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  out->clear();
  uint8_t tmp[1 << 16];
  size_t n;
  while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) out->insert(out->end(), tmp, tmp + n);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

// Writes `data` to `path` atomically: `path.tmp`, then rename.
bool WriteAtomic(const std::string& path, const void* data, size_t n) {
  const std::string tmp = path + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "wb");
  if (!f) return false;
  const bool ok = std::fwrite(data, 1, n, f) == n;
  if (std::fclose(f) != 0 || !ok) {
    std::remove(tmp.c_str());
    return false;
  }
  return std::rename(tmp.c_str(), path.c_str()) == 0;
}

// The manifest: {"items": [{"path": ..., "bytes": ..., "sha256": ...,
// "stage": ...}, ...]}.  Items are kept as their JSON text, so entries other
// writers made (other stages, other fields) survive a rewrite unchanged.
struct Item {
  std::string path, sha, text;
};

// Value of string field `key` in the JSON object text `obj` ("" if absent).
std::string StringField(const std::string& obj, const char* key) {
  const std::string k = std::string("\"") + key + "\"";
  size_t p = 0;
  while ((p = obj.find(k, p)) != std::string::npos) {
    size_t q = p + k.size();
    while (q < obj.size() && std::isspace(static_cast<unsigned char>(obj[q]))) ++q;
    if (q < obj.size() && obj[q] == ':') {
      ++q;
      while (q < obj.size() && std::isspace(static_cast<unsigned char>(obj[q]))) ++q;
      if (q >= obj.size() || obj[q] != '"') return "";
      std::string v;
      for (++q; q < obj.size() && obj[q] != '"'; ++q) {
        if (obj[q] == '\\' && q + 1 < obj.size()) ++q;
        v += obj[q];
      }
      return v;
    }
    p = q;
  }
  return "";
}

std::vector<Item> ReadManifest(const std::string& file) {
  std::vector<Item> items;
  std::vector<uint8_t> raw;
  if (!ReadFile(file, &raw)) return items;
  const std::string s(raw.begin(), raw.end());
  size_t p = s.find("\"items\"");
  if (p == std::string::npos || (p = s.find('[', p)) == std::string::npos) return items;
  // top-level objects of the array, strings (and their escapes) skipped
  int depth = 0;
  bool in_str = false;
  size_t start = 0;
  for (size_t i = p + 1; i < s.size(); ++i) {
    const char c = s[i];
    if (in_str) {
      if (c == '\\') ++i;
      else if (c == '"') in_str = false;
      continue;
    }
    if (c == '"') in_str = true;
    else if (c == '{') {
      if (depth++ == 0) start = i;
    } else if (c == '}') {
      if (--depth == 0) {
        Item it;
        it.text = s.substr(start, i - start + 1);
        it.path = StringField(it.text, "path");
        it.sha = StringField(it.text, "sha256");
        items.push_back(it);
      }
    } else if (c == ']' && depth == 0) {
      break;
    }
  }
  return items;
}

bool WriteManifest(const std::string& file, const std::vector<Item>& items) {
  std::string s = "{\"items\": [\n";
  for (size_t i = 0; i < items.size(); ++i) s += "  " + items[i].text + (i + 1 < items.size() ? ",\n" : "\n");
  s += "]}\n";
  return WriteAtomic(file, s.data(), s.size());
}

}  // namespace

extern "C" {

RECON_MVS_API int32_t recon_mvs_spill_depth_map(const char* cache_dir, int32_t ref_index, int32_t width,
                                                int32_t height, const float* depth,
                                                const uint8_t* support, const float* confidence,
                                                char* sha256_hex) {
  if (!cache_dir || ref_index < 0 || ref_index > 99999 || width <= 0 || height <= 0 || !depth ||
      !support || !confidence)
    return RECON_MVS_ERR_INVALID_ARGUMENT;
  const std::string root(cache_dir);
  if (!MakeDir(root) || !MakeDir(root + "/mvs_depth")) return RECON_MVS_ERR_INTERNAL;
  const size_t P = static_cast<size_t>(width) * height;
  std::vector<uint8_t> blob(sizeof(Header) + 3 * P * sizeof(float));
  Header hd;
  std::memset(&hd, 0, sizeof(hd));
  std::memcpy(hd.magic, "RMVD", 4);
  hd.version = 1;
  hd.width = width;
  hd.height = height;
  hd.ref_index = ref_index;
  hd.planes = 3;
  std::memcpy(blob.data(), &hd, sizeof(hd));
  float* planes = reinterpret_cast<float*>(blob.data() + sizeof(Header));
  std::memcpy(planes, depth, P * sizeof(float));
  std::memcpy(planes + P, confidence, P * sizeof(float));
  for (size_t i = 0; i < P; ++i) planes[2 * P + i] = static_cast<float>(support[i]);
  const std::string rel = MapName(ref_index);
  if (!WriteAtomic(root + "/" + rel, blob.data(), blob.size())) return RECON_MVS_ERR_INTERNAL;
  Sha256 sh;
  sh.Update(blob.data(), blob.size());
  const std::string hex = sh.HexDigest();
  std::vector<Item> items = ReadManifest(root + "/cache_manifest.json");
  Item it;
  it.path = rel;
  it.sha = hex;
  it.text = "{\"path\": \"" + rel + "\", \"bytes\": " + std::to_string(blob.size()) + ", \"sha256\": \"" +
            hex + "\", \"stage\": \"mvs\"}";
  bool replaced = false;
  for (auto& x : items)
    if (x.path == rel) {
      x = it;
      replaced = true;
    }
  if (!replaced) items.push_back(it);
  if (!WriteManifest(root + "/cache_manifest.json", items)) return RECON_MVS_ERR_INTERNAL;
  if (sha256_hex) std::memcpy(sha256_hex, hex.c_str(), 65);
  return RECON_MVS_OK;
}

RECON_MVS_API int32_t recon_mvs_load_depth_map(const char* cache_dir, int32_t ref_index, int64_t capacity,
                                               int32_t* width, int32_t* height, float* depth,
                                               uint8_t* support, float* confidence) {
  if (!cache_dir || ref_index < 0 || ref_index > 99999 || !width || !height || capacity < 0)
    return RECON_MVS_ERR_INVALID_ARGUMENT;
  const std::string root(cache_dir), rel = MapName(ref_index), path = root + "/" + rel;
  const std::string manifest = root + "/cache_manifest.json";
  std::vector<uint8_t> blob;
  if (!ReadFile(path, &blob)) return RECON_MVS_ERR_NO_DEPTH;
  std::vector<Item> items = ReadManifest(manifest);
  std::string want;
  for (const auto& it : items)
    if (it.path == rel) want = it.sha;
  Sha256 sh;
  sh.Update(blob.data(), blob.size());
  Header hd;
  std::memset(&hd, 0, sizeof(hd));
  if (blob.size() >= sizeof(hd)) std::memcpy(&hd, blob.data(), sizeof(hd));
  const size_t P = static_cast<size_t>(std::max(hd.width, 0)) * static_cast<size_t>(std::max(hd.height, 0));
  bool ok = !want.empty() && sh.HexDigest() == want && blob.size() >= sizeof(hd) &&
            std::memcmp(hd.magic, "RMVD", 4) == 0 && hd.version == 1 && hd.planes == 3 &&
            hd.ref_index == ref_index && hd.width > 0 && hd.height > 0 &&
            blob.size() == sizeof(hd) + 3 * P * sizeof(float);
  const float* planes = reinterpret_cast<const float*>(blob.data() + sizeof(Header));
  for (size_t i = 0; ok && i < P; ++i) {
    const float c = planes[2 * P + i];
    ok = c >= 0.0f && c <= 255.0f && c == std::floor(c);
  }
  if (!ok) {
    // an item that fails verification is deleted, file and manifest entry,
    // and the map is to be recomputed
    std::remove(path.c_str());
    std::vector<Item> keep;
    for (auto& it : items)
      if (it.path != rel) keep.push_back(it);
    if (keep.size() != items.size()) WriteManifest(manifest, keep);
    return RECON_MVS_ERR_NO_DEPTH;
  }
  *width = hd.width;
  *height = hd.height;
  if (!depth && !support && !confidence) return RECON_MVS_OK;  // size query
  if (capacity < static_cast<int64_t>(P)) return RECON_MVS_ERR_INVALID_ARGUMENT;
  if (depth) std::memcpy(depth, planes, P * sizeof(float));
  if (confidence) std::memcpy(confidence, planes + P, P * sizeof(float));
  if (support)
    for (size_t i = 0; i < P; ++i) support[i] = static_cast<uint8_t>(planes[2 * P + i]);
  return RECON_MVS_OK;
}

}  // extern "C"
