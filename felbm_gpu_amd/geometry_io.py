"""Geometry ingest in the reference's on-disk format: a `3d_image` TIFF stack.

Format (scripts/geometry/README.md:77-88): one TIFF page per z-slice, row = y,
column = x, so voxel (x, y, z) is element [z][y][x]; 8-bit, one sample per
pixel, strips (not tiles); grain = 255, pore = 0.  Every file with the
configured extension in `image_dir` is read in name order, each contributing
its pages as consecutive z-slices.

Domain keys (cfg/3d_image/domain.cfg:19-42): image_dir, extension,
interface_solid, invert_image, coarsening_levels, use_subvolume, x_min, y_min,
z_min, subvolume_size_x/y/z.  The reader that gives them meaning
(DomainInitializer_3DImage) is in felbm_local and not available; the
semantics used here, from the comments of that file, are:
  - interface_solid = true: a non-zero voxel is solid; false: a zero voxel is
    solid; invert_image flips the result;
  - the subvolume is cut first (full-resolution coordinates), then each
    coarsening level maps 2x2x2 voxels to one, solid when at least 4 of the 8
    are solid (odd trailing layers are dropped);
  - add_inlet_pores / add_outlet_pores are not supported (an error is raised
    when set).
Returns the solid mask uint8 [nz][ny][nx] the engine's geometry takes.

TIFF decoding follows the TIFF 6.0 specification: baseline IFD chain, strips,
compression none (1), LZW (5), Deflate (8, 32946) and PackBits (32773),
horizontal predictor (2) for LZW/Deflate; little- and big-endian files.
"""
import os
import re
import struct
import zlib

import numpy as np

_TYPES = {1: ("B", 1), 2: ("c", 1), 3: ("H", 2), 4: ("I", 4), 5: ("II", 8), 6: ("b", 1), 7: ("B", 1),
          8: ("h", 2), 9: ("i", 4), 10: ("ii", 8), 11: ("f", 4), 12: ("d", 8), 16: ("Q", 8)}


def _packbits_decode(data, expected):
    out = bytearray()
    i, n = 0, len(data)
    while i < n and len(out) < expected:
        h = data[i]
        i += 1
        if h < 128:                       # literal run of h+1 bytes
            out += data[i:i + h + 1]
            i += h + 1
        elif h > 128:                     # replicate next byte 257-h times
            out += bytes([data[i]]) * (257 - h)
            i += 1
        # h == 128: no-op
    return bytes(out[:expected])


def _lzw_decode(data, expected):
    """TIFF LZW (MSB-first codes, ClearCode 256, EOI 257, early change)."""
    out = bytearray()
    table = [bytes([i]) for i in range(256)] + [b"", b""]
    bits, nbits, pos = 0, 0, 0
    width = 9
    prev = None
    n = len(data)
    while len(out) < expected:
        while nbits < width and pos < n:
            bits = (bits << 8) | data[pos]
            pos += 1
            nbits += 8
        if nbits < width:
            break
        code = (bits >> (nbits - width)) & ((1 << width) - 1)
        nbits -= width
        if code == 257:
            break
        if code == 256:
            table = table[:258]
            width = 9
            prev = None
            continue
        if prev is None:
            entry = table[code]
            out += entry
            prev = entry
            continue
        if code < len(table):
            entry = table[code]
            table.append(prev + entry[:1])
        else:
            entry = prev + prev[:1]
            table.append(entry)
        out += entry
        prev = entry
        # early change: widen one code before the table fills the current width
        if len(table) + 1 >= (1 << width) and width < 12:
            width += 1
    return bytes(out[:expected])


class TiffError(ValueError):
    pass


def _read_ifds(buf):
    if buf[:2] == b"II":
        e = "<"
    elif buf[:2] == b"MM":
        e = ">"
    else:
        raise TiffError("not a TIFF file")
    magic, off = struct.unpack(e + "HI", buf[2:8])
    if magic != 42:
        raise TiffError("unsupported TIFF variant (BigTIFF?)")
    pages = []
    seen = set()
    while off:
        if off in seen or off + 2 > len(buf):
            raise TiffError("corrupt IFD chain")
        seen.add(off)
        (count,) = struct.unpack(e + "H", buf[off:off + 2])
        tags = {}
        for t in range(count):
            p = off + 2 + 12 * t
            tag, typ, cnt = struct.unpack(e + "HHI", buf[p:p + 8])
            if typ not in _TYPES:
                continue
            fmt, size = _TYPES[typ]
            total = size * cnt
            raw = buf[p + 8:p + 12] if total <= 4 else buf[struct.unpack(e + "I", buf[p + 8:p + 12])[0]:][:total]
            if typ in (5, 10):
                vals = struct.unpack(e + fmt[0] * (2 * cnt), raw[:total])
            elif typ == 2:
                vals = raw[:total]
            else:
                vals = struct.unpack(e + fmt * cnt, raw[:total])
            tags[tag] = vals
        pages.append(tags)
        (off,) = struct.unpack(e + "I", buf[off + 2 + 12 * count:off + 6 + 12 * count])
    return pages


def read_tiff(path):
    """All pages of one TIFF file as uint8 [pages][rows][cols]."""
    with open(path, "rb") as f:
        buf = f.read()
    out = []
    for tags in _read_ifds(buf):
        w, h = tags[256][0], tags[257][0]
        bps = tags.get(258, (1,))
        spp = tags.get(277, (1,))[0]
        if spp != 1 or any(b != 8 for b in bps):
            raise TiffError("only 8-bit single-sample images are supported")
        if 322 in tags or 324 in tags:
            raise TiffError("tiled TIFFs are not supported (the reference reads scanlines)")
        comp = tags.get(259, (1,))[0]
        pred = tags.get(317, (1,))[0]
        rps = tags.get(278, (h,))[0]
        offs, cnts = tags[273], tags[279]
        rows = []
        for s, (o, c) in enumerate(zip(offs, cnts)):
            nrow = min(rps, h - s * rps)
            expected = nrow * w
            data = buf[o:o + c]
            if comp == 1:
                raw = data[:expected]
            elif comp == 32773:
                raw = _packbits_decode(data, expected)
            elif comp == 5:
                raw = _lzw_decode(data, expected)
            elif comp in (8, 32946):
                raw = zlib.decompress(data)[:expected]
            else:
                raise TiffError("unsupported compression %d" % comp)
            a = np.frombuffer(raw, dtype=np.uint8)
            if a.size != expected:
                raise TiffError("short strip")
            a = a.reshape(nrow, w)
            if pred == 2:
                a = np.cumsum(a, axis=1, dtype=np.uint8)
            rows.append(a)
        img = np.concatenate(rows, axis=0)
        if tags.get(262, (1,))[0] == 0:          # WhiteIsZero
            img = 255 - img
        out.append(img)
    return np.stack(out, axis=0)


def read_stack(image_dir, extension="tif"):
    """Every *.extension file of a folder, in name order, pages as z-slices."""
    ext = "." + extension.lstrip(".").lower()
    names = sorted(f for f in os.listdir(image_dir) if f.lower().endswith(ext))
    if not names:
        raise TiffError("no *%s files in %s" % (ext, image_dir))
    vols = [read_tiff(os.path.join(image_dir, f)) for f in names]
    shapes = {v.shape[1:] for v in vols}
    if len(shapes) != 1:
        raise TiffError("pages of different sizes in %s" % image_dir)
    return np.concatenate(vols, axis=0)


def write_tiff(path, vol):
    """Uncompressed little-endian multi-page 8-bit TIFF, one page per [z]."""
    vol = np.ascontiguousarray(vol, dtype=np.uint8)
    nz, ny, nx = vol.shape
    parts = [b"II*\x00" + struct.pack("<I", 8)]
    pos = 8
    for z in range(nz):
        data = vol[z].tobytes()
        ntag = 9
        ifd_size = 2 + 12 * ntag + 4
        data_off = pos + ifd_size
        next_off = data_off + len(data) if z + 1 < nz else 0
        tags = [(256, 4, nx), (257, 4, ny), (258, 3, 8), (259, 3, 1), (262, 3, 1), (273, 4, data_off),
                (277, 3, 1), (278, 4, ny), (279, 4, len(data))]
        ifd = struct.pack("<H", ntag)
        for tag, typ, val in tags:
            ifd += struct.pack("<HHI", tag, typ, 1) + (struct.pack("<HH", val, 0) if typ == 3 else struct.pack("<I", val))
        ifd += struct.pack("<I", next_off)
        parts += [ifd, data]
        pos = data_off + len(data)
    with open(path, "wb") as f:
        f.write(b"".join(parts))


def parse_cfg(path):
    """key = value lines; '#' starts a comment."""
    out = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            m = re.match(r"^([A-Za-z_][A-Za-z0-9_]*)\s*=\s*(.*)$", line)
            if m:
                out[m.group(1)] = m.group(2).strip()
    return out


def _flag(v):
    return str(v).strip().lower() in ("1", "true", "yes", "on")


def coarsen(solid, levels):
    for _ in range(int(levels)):
        nz, ny, nx = (d // 2 * 2 for d in solid.shape)
        s = solid[:nz, :ny, :nx].astype(np.int32)
        s = s.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(1, 3, 5))
        solid = (s >= 4).astype(np.uint8)
    return solid


def load_3d_image(cfg, base_dir=None):
    """Solid mask [nz][ny][nx] from a domain cfg (path or dict of its keys)."""
    if isinstance(cfg, str):
        base_dir = base_dir or os.path.dirname(os.path.abspath(cfg))
        cfg = parse_cfg(cfg)
    base_dir = base_dir or "."
    if _flag(cfg.get("add_inlet_pores", "false")) or _flag(cfg.get("add_outlet_pores", "false")):
        raise NotImplementedError("add_inlet_pores / add_outlet_pores are not supported")
    image_dir = cfg["image_dir"]
    if not os.path.isabs(image_dir):
        image_dir = os.path.join(base_dir, image_dir)
    vol = read_stack(image_dir, cfg.get("extension", "tif"))
    if _flag(cfg.get("use_subvolume", "false")):
        x0, y0, z0 = (int(cfg.get(k, 0)) for k in ("x_min", "y_min", "z_min"))
        sx, sy, sz = (int(cfg[k]) for k in ("subvolume_size_x", "subvolume_size_y", "subvolume_size_z"))
        vol = vol[z0:z0 + sz, y0:y0 + sy, x0:x0 + sx]
    nonzero_solid = _flag(cfg.get("interface_solid", "true")) != _flag(cfg.get("invert_image", "false"))
    solid = ((vol != 0) == nonzero_solid).astype(np.uint8)
    return coarsen(solid, int(cfg.get("coarsening_levels", 0)))
