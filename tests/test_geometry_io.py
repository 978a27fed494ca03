"""Geometry ingest from `3d_image` TIFF stacks (felbm_gpu_amd/geometry_io.py).

Pins: the PackBits known-answer example of the TIFF 6.0 specification (section
9); independent encodings written by Pillow (uncompressed, PackBits, LZW,
Deflate, multi-page) decode to the same voxels; a big-endian two-strip file
built byte by byte; the voxel convention (page = z, row = y, column = x); and
the domain keys (interface_solid / invert_image, subvolume, coarsening by
2x2x2 majority) on hand-checkable volumes.
"""
import os
import struct

import numpy as np
import pytest

from felbm_gpu_amd import geometry_io as G


def test_packbits_known_answer_from_tiff6_spec():
    packed = bytes.fromhex("FE AA 02 80 00 2A FD AA 03 80 00 2A 22 F7 AA".replace(" ", ""))
    expect = bytes.fromhex("AA AA AA 80 00 2A AA AA AA AA 80 00 2A 22 AA AA AA AA AA AA AA AA AA AA".replace(" ", ""))
    assert G._packbits_decode(packed, len(expect)) == expect


def test_write_read_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    vol = (rng.random((5, 7, 9)) < 0.4).astype(np.uint8) * 255
    p = tmp_path / "a.tif"
    G.write_tiff(str(p), vol)
    assert np.array_equal(G.read_tiff(str(p)), vol)


@pytest.mark.parametrize("compression", [None, "packbits", "tiff_lzw", "tiff_deflate"])
def test_pillow_encodings_decode_identically(tmp_path, compression):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(2)
    vol = (rng.random((4, 37, 53)) < 0.5).astype(np.uint8) * 255
    vol[1, :, :20] = 0
    vol[2, 10:, :] = 255                         # long runs for the LZW/PackBits tables
    pages = [Image.fromarray(vol[z]) for z in range(vol.shape[0])]
    p = tmp_path / "b.tif"
    kw = {"compression": compression} if compression else {}
    pages[0].save(str(p), save_all=True, append_images=pages[1:], **kw)
    assert np.array_equal(G.read_tiff(str(p)), vol)


def test_lzw_table_growth_on_a_large_page(tmp_path):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, (300, 300), dtype=np.uint8)      # fills the 12-bit table and clears
    p = tmp_path / "c.tif"
    Image.fromarray(img).save(str(p), compression="tiff_lzw")
    assert np.array_equal(G.read_tiff(str(p))[0], img)


def test_big_endian_two_strips_by_hand(tmp_path):
    w, h = 3, 4
    pix = np.arange(12, dtype=np.uint8).reshape(h, w) * 20
    data = pix.tobytes()
    # header + IFD at offset 8, strips after the IFD
    ntag = 8
    strip0 = 8 + 2 + 12 * ntag + 4
    strip1 = strip0 + 6
    soff = strip1 + 6                               # StripOffsets array (2 LONGs)
    scnt = soff + 8                                 # StripByteCounts array (2 LONGs)
    ifd = struct.pack(">H", ntag)
    for tag, typ, cnt, val in [(256, 3, 1, w), (257, 3, 1, h), (258, 3, 1, 8), (259, 3, 1, 1), (262, 3, 1, 1),
                               (273, 4, 2, soff), (278, 3, 1, 2), (279, 4, 2, scnt)]:
        v = struct.pack(">HH", val, 0) if (typ == 3 and cnt == 1) else struct.pack(">I", val)
        ifd += struct.pack(">HHI", tag, typ, cnt) + v
    ifd += struct.pack(">I", 0)
    buf = b"MM\x00\x2a" + struct.pack(">I", 8) + ifd + data[:6] + data[6:] + struct.pack(">II", strip0, strip1) \
        + struct.pack(">II", 6, 6)
    p = tmp_path / "d.tif"
    p.write_bytes(buf)
    assert np.array_equal(G.read_tiff(str(p))[0], pix)


def test_voxel_convention_and_stack_order(tmp_path):
    d = tmp_path / "img"
    d.mkdir()
    a = np.zeros((2, 3, 4), np.uint8)
    a[0, 1, 2] = 255                                 # voxel (x=2, y=1, z=0)
    b = np.zeros((1, 3, 4), np.uint8)
    b[0, 2, 3] = 255                                 # voxel (x=3, y=2, z=2) after stacking
    G.write_tiff(str(d / "s_000.tif"), a)
    G.write_tiff(str(d / "s_001.tif"), b)
    (d / "notes.txt").write_text("ignored")
    vol = G.read_stack(str(d), "tif")
    assert vol.shape == (3, 3, 4)
    assert vol[0, 1, 2] == 255 and vol[2, 2, 3] == 255 and vol.sum() == 2 * 255


def _cfg(tmp_path, vol, **keys):
    d = tmp_path / "image"
    d.mkdir(exist_ok=True)
    G.write_tiff(str(d / "stack.tif"), vol)
    base = {"image_dir": "./image/", "extension": "tif", "interface_solid": "true", "invert_image": "false",
            "coarsening_levels": "0", "use_subvolume": "false", "x_min": "0", "y_min": "0", "z_min": "0",
            "subvolume_size_x": "1", "subvolume_size_y": "1", "subvolume_size_z": "1",
            "add_inlet_pores": "false", "add_outlet_pores": "false"}
    base.update({k: str(v) for k, v in keys.items()})
    path = tmp_path / "domain.cfg"
    path.write_text("# test domain\n" + "".join("%s = %s   # comment\n" % kv for kv in base.items()))
    return str(path)


def test_domain_keys(tmp_path):
    rng = np.random.default_rng(4)
    vol = (rng.random((8, 10, 12)) < 0.3).astype(np.uint8) * 255
    cfg = _cfg(tmp_path, vol)
    assert np.array_equal(G.load_3d_image(cfg), (vol != 0).astype(np.uint8))
    cfg = _cfg(tmp_path, vol, invert_image="true")
    assert np.array_equal(G.load_3d_image(cfg), (vol == 0).astype(np.uint8))
    cfg = _cfg(tmp_path, vol, interface_solid="false")
    assert np.array_equal(G.load_3d_image(cfg), (vol == 0).astype(np.uint8))
    cfg = _cfg(tmp_path, vol, use_subvolume="true", x_min=2, y_min=3, z_min=1, subvolume_size_x=5,
               subvolume_size_y=4, subvolume_size_z=6)
    assert np.array_equal(G.load_3d_image(cfg), (vol[1:7, 3:7, 2:7] != 0).astype(np.uint8))
    with pytest.raises(NotImplementedError):
        G.load_3d_image(_cfg(tmp_path, vol, add_inlet_pores="true"))


def test_coarsening_is_a_2x2x2_majority():
    s = np.zeros((2, 2, 4), np.uint8)
    s[:, :, :2].flat[:3] = 1                         # 3 of 8 solid -> fluid
    s[:, :, 2:].flat[:0] = 1
    s[0, 0, 2] = s[0, 1, 2] = s[1, 0, 3] = s[1, 1, 3] = 1   # 4 of 8 -> solid
    c = G.coarsen(s, 1)
    assert c.shape == (1, 1, 2) and c[0, 0, 0] == 0 and c[0, 0, 1] == 1
    odd = np.ones((5, 5, 5), np.uint8)
    assert G.coarsen(odd, 1).shape == (2, 2, 2)
    assert G.coarsen(np.ones((8, 8, 8), np.uint8), 2).shape == (2, 2, 2)


def test_ingested_geometry_runs_in_the_oracle(tmp_path):
    from felbm_gpu_amd import workloads as W
    from oracle.pyoracle import Oracle
    grains = W.spheres(16, 16, 16, 2.0, 3.0, 0.6, 5, keep_largest=False) * 255
    cfg = _cfg(tmp_path, grains)
    solid = G.load_3d_image(cfg)
    assert np.array_equal(solid, (grains != 0).astype(np.uint8))
    o = Oracle(solid, W.derive_params(), use_mrt=True)
    o.init_rest(W.slab_c(solid, 0, 0, 8, 4.0), 1.0)
    o.run(2)
    assert np.isfinite(o.download()["c"]).all()
