"""Scene files (CPU): the versioned format round-trips byte for byte, the reference's raw
.bvhcache dumps import, and damaged or malformed files are refused with a clear status."""
import numpy as np
import pytest

from phantasyenginetestbed_amd import phx


@pytest.fixture(scope="module")
def scene():
    tris, datas = phx.heightfield_boxes(24, 10, 0xC0FFEE02, num_materials=8)
    bvh = phx.BVH(tris, datas)
    return bvh, phx.materials_random(8, 1), phx.lights_random(3, 1)


def test_round_trip_is_byte_identical(tmp_path, scene):
    bvh, mats, lights = scene
    p = tmp_path / "s.phxscene"
    phx.save_scene(p, bvh.nodes, bvh.tris, bvh.datas, mats, lights)
    raw = p.read_bytes()
    assert raw[:8] == b"PHXSCENE" and len(raw) % 64 == 0
    got = phx.load_scene(p)
    for k, want in (("nodes", bvh.nodes), ("tris", bvh.tris), ("datas", bvh.datas),
                    ("materials", mats), ("lights", lights)):
        assert got[k].tobytes() == want.tobytes(), k
    # without datas, materials, lights
    q = tmp_path / "bare.phxscene"
    phx.save_scene(q, bvh.nodes, bvh.tris)
    bare = phx.load_scene(q)
    assert bare["datas"] is None and len(bare["materials"]) == 0 and len(bare["lights"]) == 0
    assert bare["nodes"].tobytes() == bvh.nodes.tobytes()


def test_reference_bvhcache_import(tmp_path, scene):
    bvh, _, _ = scene
    # the reference writes each array as a raw dump (Main.cpp:149-151)
    (tmp_path / "n.bvhcache").write_bytes(bvh.nodes.tobytes())
    (tmp_path / "v.bvhcache").write_bytes(bvh.tris.tobytes())
    (tmp_path / "d.bvhcache").write_bytes(bvh.datas.tobytes())
    got = phx.load_bvhcache(tmp_path / "n.bvhcache", tmp_path / "v.bvhcache",
                            tmp_path / "d.bvhcache")
    assert got["nodes"].tobytes() == bvh.nodes.tobytes()
    assert got["tris"].tobytes() == bvh.tris.tobytes()
    assert got["datas"].tobytes() == bvh.datas.tobytes()
    # a datas dump of another length, or a partial record, is refused
    (tmp_path / "short.bvhcache").write_bytes(bvh.datas[:-1].tobytes())
    with pytest.raises(phx.PhxError) as e:
        phx.load_bvhcache(tmp_path / "n.bvhcache", tmp_path / "v.bvhcache",
                          tmp_path / "short.bvhcache")
    assert e.value.status == phx.ERR_INVALID_SCENE
    (tmp_path / "odd.bvhcache").write_bytes(bvh.nodes.tobytes()[:-5])
    with pytest.raises(phx.PhxError):
        phx.load_bvhcache(tmp_path / "odd.bvhcache", tmp_path / "v.bvhcache")


def test_damaged_files_are_refused(tmp_path, scene):
    bvh, mats, lights = scene
    p = tmp_path / "s.phxscene"
    phx.save_scene(p, bvh.nodes, bvh.tris, bvh.datas, mats, lights)
    raw = bytearray(p.read_bytes())

    def refused(data, status=phx.ERR_INVALID_SCENE):
        q = tmp_path / "bad.phxscene"
        q.write_bytes(bytes(data))
        with pytest.raises(phx.PhxError) as e:
            phx.load_scene(q)
        assert e.value.status == status, e.value
        return str(e.value)

    assert "not a phx scene" in refused(b"BVHCACHE" + raw[8:])
    bad = bytearray(raw)
    bad[8] = 2  # version
    assert "version" in refused(bad)
    assert "truncated" in refused(raw[:-100])
    assert "trailing" in refused(raw + b"\0" * 64)
    # A header whose counts the file cannot hold is refused before any section is sized from
    # it (here 2^27 triangles: 15 GB of triangle and data sections).
    big = bytearray(raw)
    big[20:24] = (1 << 27).to_bytes(4, "little")
    assert "truncated" in refused(big)
    flipped = bytearray(raw)
    flipped[64 + 5] ^= 0x40  # a bit inside the first node
    assert "checksum" in refused(flipped)
    with pytest.raises(phx.PhxError):
        phx.load_scene(tmp_path / "missing.phxscene")


def test_malformed_bvh_is_refused_on_write_and_read(tmp_path, scene):
    bvh, _, _ = scene
    nodes = bvh.nodes.copy()
    nodes["i"][0, 0] = len(nodes) + 7  # left child out of range
    with pytest.raises(phx.PhxError) as e:
        phx.save_scene(tmp_path / "w.phxscene", nodes, bvh.tris)
    assert e.value.status == phx.ERR_MALFORMED_BVH
    (tmp_path / "n.bvhcache").write_bytes(nodes.tobytes())
    (tmp_path / "v.bvhcache").write_bytes(bvh.tris.tobytes())
    with pytest.raises(phx.PhxError) as e:
        phx.load_bvhcache(tmp_path / "n.bvhcache", tmp_path / "v.bvhcache")
    assert e.value.status == phx.ERR_MALFORMED_BVH
    # a material index past the material table
    datas = bvh.datas.copy()
    datas["mat"][3] = 99
    with pytest.raises(phx.PhxError):
        phx.save_scene(tmp_path / "m.phxscene", bvh.nodes, bvh.tris, datas,
                       phx.materials_random(8, 1))
        phx.load_scene(tmp_path / "m.phxscene")
