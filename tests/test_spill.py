"""Depth-map spill to the session cache (docs/13_内存与降级.md §3.3.2-3.3.4):
`<cache>/mvs_depth/ref_%05d.f32` holding depth, confidence and support counts,
written to *.tmp then renamed, recorded in `cache_manifest.json` with path,
bytes, sha256 and stage, verified on read (a failed item is deleted and the
map reported missing, so the caller recomputes it).

The reference holds this format only as that specification, no code and no
fixtures.  The library's bytes are checked against the oracle's encoder
(numpy) in both directions, and its SHA-256 against hashlib, over map sizes
whose files end on each side of the SHA-256 padding boundaries.  Host-only
entry points: these run on the CPU."""
import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from recon_scanner_amd import _native as N
from recon_scanner_amd import mvs


def random_map(w, h, seed):
    rng = np.random.default_rng(seed)
    dm = mvs.DepthMap()
    dm.Allocate(w, h)
    dm.depth[:] = rng.uniform(0.1, 2.0, w * h).astype(np.float32)
    dm.depth[rng.random(w * h) < 0.2] = np.nan
    dm.support_count[:] = rng.integers(0, 256, w * h, dtype=np.uint8)
    dm.confidence[:] = rng.random(w * h).astype(np.float32)
    return dm


def manifest(cache):
    with open(os.path.join(cache, "cache_manifest.json")) as f:
        return json.load(f)["items"]


# (w, h): file sizes 44, 56, 68, 152, 32 + 12 * 4800 bytes: one, two and
# many SHA-256 blocks, the length field landing in the first / a second block
@pytest.mark.parametrize("w,h", [(1, 1), (2, 1), (3, 1), (5, 2), (80, 60)])
def test_spill_bytes_and_digest(tmp_path, w, h):
    dm = random_map(w, h, seed=w * 100 + h)
    sha = mvs.SpillDepthMap(str(tmp_path), 7, dm)
    path = tmp_path / "mvs_depth" / "ref_00007.f32"
    blob = path.read_bytes()
    assert blob == oracle.spill_encode(7, w, h, dm.depth, dm.support_count, dm.confidence)
    assert sha == hashlib.sha256(blob).hexdigest()
    assert manifest(str(tmp_path)) == [oracle.spill_manifest_item(7, blob)]
    assert not any(p.name.endswith(".tmp") for p in tmp_path.rglob("*"))


def test_round_trip_bit_exact(tmp_path):
    dm = random_map(64, 48, seed=3)
    mvs.SpillDepthMap(str(tmp_path), 12, dm)
    back = mvs.LoadDepthMap(str(tmp_path), 12)
    assert (back.width, back.height) == (64, 48)
    assert back.depth.tobytes() == dm.depth.tobytes()  # NaNs included
    assert np.array_equal(back.support_count, dm.support_count)
    assert back.confidence.tobytes() == dm.confidence.tobytes()


def test_oracle_written_map_is_read(tmp_path):
    """A map written by the oracle's encoder with a manifest written by Python's
    json (other spacing, other stages' items, escapes) reads back."""
    dm = random_map(9, 4, seed=5)
    blob = oracle.spill_encode(3, 9, 4, dm.depth, dm.support_count, dm.confidence)
    (tmp_path / "mvs_depth").mkdir()
    (tmp_path / "mvs_depth" / "ref_00003.f32").write_bytes(blob)
    other = {"path": "tsdf_blocks/block_00001.bin", "bytes": 10, "sha256": "0" * 64,
             "stage": "tsdf", "note": "braces } { and \"quotes\""}
    items = [other, oracle.spill_manifest_item(3, blob)]
    (tmp_path / "cache_manifest.json").write_text(json.dumps({"items": items}, indent=4))
    back = mvs.LoadDepthMap(str(tmp_path), 3)
    assert back.depth.tobytes() == dm.depth.tobytes()
    assert np.array_equal(back.support_count, dm.support_count)
    # a later spill keeps the other writer's item unchanged and replaces its own
    mvs.SpillDepthMap(str(tmp_path), 3, random_map(9, 4, seed=6))
    got = manifest(str(tmp_path))
    assert got[0] == other
    assert [it["path"] for it in got] == [other["path"], "mvs_depth/ref_00003.f32"]


def test_respill_replaces_and_maps_accumulate(tmp_path):
    a, b, c = random_map(8, 8, 1), random_map(8, 8, 2), random_map(4, 4, 3)
    mvs.SpillDepthMap(str(tmp_path), 0, a)
    mvs.SpillDepthMap(str(tmp_path), 1, c)
    mvs.SpillDepthMap(str(tmp_path), 0, b)
    items = manifest(str(tmp_path))
    assert sorted(it["path"] for it in items) == ["mvs_depth/ref_00000.f32", "mvs_depth/ref_00001.f32"]
    assert mvs.LoadDepthMap(str(tmp_path), 0).depth.tobytes() == b.depth.tobytes()
    assert mvs.LoadDepthMap(str(tmp_path), 1).confidence.tobytes() == c.confidence.tobytes()


@pytest.mark.parametrize("damage", ["flip_payload", "flip_header", "truncate", "no_manifest_item",
                                    "bad_support"])
def test_failed_verification_deletes_the_item(tmp_path, damage):
    keep = random_map(4, 4, 9)
    mvs.SpillDepthMap(str(tmp_path), 2, keep)
    dm = random_map(16, 8, 4)
    mvs.SpillDepthMap(str(tmp_path), 5, dm)
    path = tmp_path / "mvs_depth" / "ref_00005.f32"
    blob = bytearray(path.read_bytes())
    if damage == "flip_payload":
        blob[100] ^= 0x01
        path.write_bytes(bytes(blob))
    elif damage == "flip_header":
        blob[8] ^= 0x01
        path.write_bytes(bytes(blob))
    elif damage == "truncate":
        path.write_bytes(bytes(blob[:-4]))
    elif damage == "no_manifest_item":
        items = [it for it in manifest(str(tmp_path)) if it["path"] != "mvs_depth/ref_00005.f32"]
        (tmp_path / "cache_manifest.json").write_text(json.dumps({"items": items}))
    else:
        # a support value out of 0..255 with a matching digest: well-formed
        # bytes are required, not only a matching digest
        f = np.frombuffer(bytes(blob), np.float32, offset=32).copy()
        f[2 * 128] = 300.0
        blob = bytes(blob[:32]) + f.tobytes()
        path.write_bytes(blob)
        items = manifest(str(tmp_path))
        for it in items:
            if it["path"] == "mvs_depth/ref_00005.f32":
                it["sha256"] = hashlib.sha256(blob).hexdigest()
        (tmp_path / "cache_manifest.json").write_text(json.dumps({"items": items}))
    assert mvs.LoadDepthMap(str(tmp_path), 5) is None
    assert not path.exists()
    assert [it["path"] for it in manifest(str(tmp_path))] == ["mvs_depth/ref_00002.f32"]
    assert mvs.LoadDepthMap(str(tmp_path), 2).depth.tobytes() == keep.depth.tobytes()


def test_missing_and_invalid_arguments(tmp_path):
    assert mvs.LoadDepthMap(str(tmp_path), 1) is None
    L = mvs.lib()
    d = np.zeros(4, np.float32)
    s = np.zeros(4, np.uint8)
    cache = os.fsencode(str(tmp_path))
    assert L.recon_mvs_spill_depth_map(cache, 100000, 2, 2, d.ctypes.data, s.ctypes.data,
                                       d.ctypes.data, None) == N.ERR_INVALID_ARGUMENT
    assert L.recon_mvs_spill_depth_map(cache, 1, 0, 2, d.ctypes.data, s.ctypes.data,
                                       d.ctypes.data, None) == N.ERR_INVALID_ARGUMENT
    assert L.recon_mvs_spill_depth_map(cache, 1, 2, 2, None, s.ctypes.data, d.ctypes.data,
                                       None) == N.ERR_INVALID_ARGUMENT
    assert L.recon_mvs_spill_depth_map(cache, 1, 2, 2, d.ctypes.data, s.ctypes.data,
                                       d.ctypes.data, None) == N.OK
    import ctypes as C
    w, h = C.c_int32(0), C.c_int32(0)
    # a buffer smaller than the map is refused, not overrun
    assert L.recon_mvs_load_depth_map(cache, 1, 3, C.byref(w), C.byref(h), d.ctypes.data, None,
                                      None) == N.ERR_INVALID_ARGUMENT
    assert (w.value, h.value) == (2, 2)
    with pytest.raises(ValueError):
        bad = mvs.DepthMap()
        bad.Allocate(3, 3)
        bad.depth = bad.depth[:4]
        mvs.SpillDepthMap(str(tmp_path), 1, bad)


def test_oriented_cloud_restatement():
    """UnifiedPointCloud::ToOrientedCloud's two rules (orchestrator.cpp:186-191):
    negative confidences weigh 0, points without a confidence weigh 1."""
    pos = np.arange(12, dtype=np.float64).reshape(4, 3)
    _, _, _, w = oracle.to_oriented_cloud(pos, pos, np.zeros(12, np.float32),
                                          np.array([0.5, -0.25, 2.0], np.float32))
    assert w.tolist() == [0.5, 0.0, 2.0, 1.0]
