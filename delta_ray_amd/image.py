"""Image files (reference image.py:8-49): ``imwrite(img, filename)`` and
``imread(filename)`` on float32 [H, W, 3] arrays.

* ``.exr``: written as three HALF channels (R, G, B) with ZIP compression (16
  scanlines per zlib block), which is what the reference's OpenEXR call
  produces by default; read back as float32.  The reader takes scanline
  files with NONE, RLE, ZIPS or ZIP compression and UINT / HALF / FLOAT
  channels (the reference's scenes ship ZIP-compressed HALF files).  The
  OpenEXR format itself is restated here from its published file layout
  (magic 20000630, attribute header, offset table, scanline blocks; ZIP and
  RLE blocks reorder bytes into two halves and store byte differences).
* ``.hdr``: Radiance RGBE (read only, as the reference, which reads it with
  imageio), flat or run-length encoded scanlines.
* ``.pfm``: portable float map (not in the reference; lossless float32).
* anything else: 8-bit through PIL -- written as clip(img, 0, 1)^(1/2.2),
  read as (value / 255)^2.2, grey images repeated to three channels, alpha
  dropped, as the reference does with scikit-image.
"""
import struct
import zlib

import numpy as np

_EXR_MAGIC = 20000630
_NONE, _RLE, _ZIPS, _ZIP = 0, 1, 2, 3
_LINES = {_NONE: 1, _RLE: 1, _ZIPS: 1, _ZIP: 16}
_PTYPE = {0: np.dtype("<u4"), 1: np.dtype("<f2"), 2: np.dtype("<f4")}


# ------------------------------------------------------------------ EXR ----
def _attr(name, typ, data):
    return name.encode() + b"\0" + typ.encode() + b"\0" + struct.pack("<i", len(data)) + data


def _reorder(raw):
    """bytes -> the two-halves order with byte differences (ZIP / RLE)"""
    b = np.frombuffer(raw, np.uint8)
    t = np.concatenate([b[0::2], b[1::2]]).astype(np.int16)
    d = t.copy()
    d[1:] = (t[1:] - t[:-1] + 128) & 0xFF
    return d.astype(np.uint8).tobytes()


def _unreorder(data, n):
    t = np.frombuffer(data, np.uint8).astype(np.int64)
    if len(t) != n:
        raise ValueError("EXR block: %d bytes after decompression, %d expected" % (len(t), n))
    t[1:] -= 128
    t = (np.cumsum(t) & 0xFF).astype(np.uint8)
    out = np.empty(n, np.uint8)
    h = (n + 1) // 2
    out[0::2] = t[:h]
    out[1::2] = t[h:]
    return out.tobytes()


def _rle_decode(data):
    out = bytearray()
    i = 0
    while i < len(data):
        c = struct.unpack("b", data[i:i + 1])[0]
        i += 1
        if c < 0:
            out += data[i:i - c]
            i -= c
        else:
            out += data[i:i + 1] * (c + 1)
            i += 1
    return bytes(out)


def write_exr(img, filename):
    img = np.asarray(img, np.float32)
    H, W = img.shape[:2]
    names = ["B", "G", "R"]  # channel lists are sorted by name
    chans = {"R": img[:, :, 0], "G": img[:, :, 1], "B": img[:, :, 2]}
    chlist = b"".join(n.encode() + b"\0" + struct.pack("<iB3xii", 1, 0, 1, 1) for n in names) + b"\0"
    header = struct.pack("<ii", _EXR_MAGIC, 2)
    header += _attr("channels", "chlist", chlist)
    header += _attr("compression", "compression", bytes([_ZIP]))
    header += _attr("dataWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
    header += _attr("displayWindow", "box2i", struct.pack("<4i", 0, 0, W - 1, H - 1))
    header += _attr("lineOrder", "lineOrder", bytes([0]))
    header += _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
    header += _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
    header += _attr("screenWindowWidth", "float", struct.pack("<f", 1.0))
    header += b"\0"
    lines = _LINES[_ZIP]
    nblocks = (H + lines - 1) // lines
    half = {n: chans[n].astype("<f2") for n in names}
    blocks = []
    for k in range(nblocks):
        y0, y1 = k * lines, min(H, (k + 1) * lines)
        raw = b"".join(half[n][y].tobytes() for y in range(y0, y1) for n in names)
        comp = zlib.compress(_reorder(raw))
        data = comp if len(comp) < len(raw) else raw
        blocks.append(struct.pack("<ii", y0, len(data)) + data)
    table_pos = len(header)
    pos = table_pos + 8 * nblocks
    offsets = []
    for b in blocks:
        offsets.append(pos)
        pos += len(b)
    with open(filename, "wb") as f:
        f.write(header)
        f.write(struct.pack("<%dQ" % nblocks, *offsets))
        for b in blocks:
            f.write(b)


def read_exr(filename):
    with open(filename, "rb") as f:
        buf = f.read()
    magic, version = struct.unpack_from("<ii", buf, 0)
    if magic != _EXR_MAGIC:
        raise ValueError("%s: not an OpenEXR file" % filename)
    if version & 0x200:
        raise ValueError("%s: tiled OpenEXR files are not supported" % filename)
    if version & 0x1800:
        raise ValueError("%s: multi-part / deep OpenEXR files are not supported" % filename)
    i = 8
    attrs = {}
    while buf[i] != 0:
        j = buf.index(b"\0", i)
        name = buf[i:j].decode()
        k = buf.index(b"\0", j + 1)
        size = struct.unpack_from("<i", buf, k + 1)[0]
        attrs[name] = buf[k + 5:k + 5 + size]
        i = k + 5 + size
    i += 1
    comp = attrs["compression"][0]
    if comp not in _LINES:
        raise ValueError("%s: OpenEXR compression %d is not supported" % (filename, comp))
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"])
    W, H = x1 - x0 + 1, y1 - y0 + 1
    cl = attrs["channels"]
    chans = []
    k = 0
    while cl[k] != 0:
        j = cl.index(b"\0", k)
        name = cl[k:j].decode()
        ptype, _, xs, ys = struct.unpack_from("<iB3xii", cl, j + 1)
        if xs != 1 or ys != 1:
            raise ValueError("%s: subsampled channels are not supported" % filename)
        chans.append((name, _PTYPE[ptype]))
        k = j + 1 + 16
    lines = _LINES[comp]
    nblocks = (H + lines - 1) // lines
    offsets = struct.unpack_from("<%dQ" % nblocks, buf, i)
    line_bytes = sum(dt.itemsize for _, dt in chans) * W
    planes = {n: np.zeros((H, W), np.float32) for n, _ in chans}
    for off in offsets:
        y, size = struct.unpack_from("<ii", buf, off)
        data = buf[off + 8:off + 8 + size]
        ly = min(lines, y1 + 1 - y)
        n = line_bytes * ly
        if size < n:
            if comp in (_ZIP, _ZIPS):
                data = _unreorder(zlib.decompress(data), n)
            elif comp == _RLE:
                data = _unreorder(_rle_decode(data), n)
        p = 0
        for r in range(ly):
            for name, dt in chans:
                planes[name][y - y0 + r] = np.frombuffer(data, dt, W, p).astype(np.float32)
                p += dt.itemsize * W
    return np.stack([planes.get(c, np.zeros((H, W), np.float32)) for c in ("R", "G", "B")], -1)


# ------------------------------------------------------------------ PFM ----
def write_pfm(img, filename):
    img = np.asarray(img, np.float32)
    H, W = img.shape[:2]
    with open(filename, "wb") as f:
        f.write(b"PF\n%d %d\n-1.0\n" % (W, H))
        f.write(np.ascontiguousarray(img[::-1, :, :3], "<f4").tobytes())


def read_pfm(filename):
    with open(filename, "rb") as f:
        kind = f.readline().strip()
        W, H = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        ch = 3 if kind == b"PF" else 1
        dt = "<f4" if scale < 0 else ">f4"
        a = np.frombuffer(f.read(4 * W * H * ch), dt).astype(np.float32).reshape(H, W, ch)[::-1]
    if ch == 1:
        a = np.repeat(a, 3, axis=2)
    return np.ascontiguousarray(a)


# ---------------------------------------------------------- Radiance HDR ----
def read_hdr(filename):
    with open(filename, "rb") as f:
        buf = f.read()
    i = 0
    while True:
        j = buf.index(b"\n", i)
        line = buf[i:j]
        i = j + 1
        if line == b"":
            break
    j = buf.index(b"\n", i)
    res = buf[i:j].split()
    i = j + 1
    if res[0] != b"-Y" or res[2] != b"+X":
        raise ValueError("%s: only -Y H +X W Radiance files are supported" % filename)
    H, W = int(res[1]), int(res[3])
    rgbe = np.zeros((H, W, 4), np.uint8)
    for y in range(H):
        if W >= 8 and W < 32768 and buf[i] == 2 and buf[i + 1] == 2 and (buf[i + 2] << 8 | buf[i + 3]) == W:
            i += 4
            for c in range(4):
                x = 0
                while x < W:
                    n = buf[i]
                    i += 1
                    if n > 128:
                        rgbe[y, x:x + n - 128, c] = buf[i]
                        i += 1
                        x += n - 128
                    else:
                        rgbe[y, x:x + n, c] = np.frombuffer(buf, np.uint8, n, i)
                        i += n
                        x += n
        else:
            rgbe[y] = np.frombuffer(buf, np.uint8, 4 * W, i).reshape(W, 4)
            i += 4 * W
    e = rgbe[:, :, 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(1.0, e - 136), 0.0).astype(np.float32)
    return rgbe[:, :, :3].astype(np.float32) * scale[:, :, None]


# -------------------------------------------------------------- front --
def imwrite(img, filename):
    img = np.asarray(img, np.float32)
    low = filename.lower()
    if low.endswith(".exr"):
        write_exr(img, filename)
    elif low.endswith(".pfm"):
        write_pfm(img, filename)
    else:
        from PIL import Image
        v = np.power(np.clip(img, 0.0, 1.0), 1.0 / 2.2)
        Image.fromarray(np.floor(v * 255.0 + 0.5).astype(np.uint8)).save(filename)


def imread(filename):
    low = filename.lower()
    if low.endswith(".exr"):
        return read_exr(filename)
    if low.endswith(".pfm"):
        return read_pfm(filename)
    if low.endswith(".hdr"):
        return read_hdr(filename)
    from PIL import Image
    im = np.asarray(Image.open(filename))
    if im.ndim == 2:
        im = np.stack([im, im, im], axis=-1)
    elif im.shape[2] == 4:
        im = im[:, :, :3]
    scale = 65535.0 if im.dtype == np.uint16 else 255.0
    return np.power(im.astype(np.float32) / np.float32(scale), np.float32(2.2))
