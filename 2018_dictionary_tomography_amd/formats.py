"""On-disk formats of the reconstruction scripts: MRC volumes and tilt
series, multi-page TIFF tilt series with `.rawtlt` angle files, and the
`savemat` outputs.

    read_mrc(path) -> (data, header)         MRC2014 reader (the scripts use
                                             odl.contrib.mrc.FileReaderMRC,
                                             polyII_recon_Newton.py:18-22)
    write_mrc(path, data, voxel_size=1.0)    MRC2014 writer (mode 2 / 1 / 0 / 6 / 12)
    read_tiff_stack(path) -> (n, rows, cols) baseline multi-page TIFF (the
                                             scripts use PIL, DogBones_recon_small.py:17-24)
    read_rawtlt(path) -> angles in radians   loadtxt(...) * pi / 180
    load_tilt_series(tif, rawtlt)            both, in the scripts' layout:
                                             data[angle, width, slice]
    save_recon(path, Radon, recon)           savemat with 'view', 'X', 'R', 'I'
                                             (Bipyramid_recon.py:140-144)

MRC layout: a 1024-byte header of 56 little- or big-endian 4-byte words and
ten 80-byte labels, an extended header of NSYMBT bytes, then the voxels with
the column index fastest.  `read_mrc` returns the voxels indexed
[column, row, section] (shape (NX, NY, NZ)): the reference's scripts index
odl's reader output that way (polyII_recon.py:28 transposes it with
(2, 1, 0) to get [angle, row, column] from a tilt series).

TIFF: uncompressed strips, one sample per pixel, 8/16/32/64-bit unsigned,
signed or IEEE samples, either byte order.  Each page is returned as
[row, column]; `load_tilt_series` transposes each page, as the scripts'
asarray(page).T does, so that the result is [angle, width, slice].
"""
from __future__ import annotations

import struct

import numpy as np

# ---------------------------------------------------------------------------
# MRC2014
# ---------------------------------------------------------------------------
_MRC_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 4: np.complex64, 6: np.uint16, 12: np.float16}
_MRC_CODES = {np.dtype(v): k for k, v in _MRC_MODES.items()}


def _mrc_endian(raw):
    """'<' or '>' from the machine stamp (word 54), falling back to a
    plausibility test of the mode word."""
    stamp = raw[212:214]
    if stamp in (b"\x44\x44", b"\x44\x41"):
        return "<"
    if stamp == b"\x11\x11":
        return ">"
    mode_le = struct.unpack("<i", raw[12:16])[0]
    return "<" if mode_le in _MRC_MODES else ">"


def read_mrc_header(raw):
    """Header fields of an MRC2014 file from its first 1024 bytes."""
    if len(raw) < 1024:
        raise ValueError("an MRC file starts with a 1024-byte header")
    e = _mrc_endian(raw)
    i = lambda k: struct.unpack(e + "i", raw[4 * k:4 * k + 4])[0]
    f = lambda k: struct.unpack(e + "f", raw[4 * k:4 * k + 4])[0]
    h = {
        "endian": e,
        "nx": i(0), "ny": i(1), "nz": i(2), "mode": i(3),
        "nxstart": i(4), "nystart": i(5), "nzstart": i(6),
        "mx": i(7), "my": i(8), "mz": i(9),
        "cella": (f(10), f(11), f(12)), "cellb": (f(13), f(14), f(15)),
        "mapc": i(16), "mapr": i(17), "maps": i(18),
        "dmin": f(19), "dmax": f(20), "dmean": f(21),
        "ispg": i(22), "nsymbt": i(23),
        "exttyp": raw[104:108], "nversion": i(27),
        "origin": (f(49), f(50), f(51)),
        "map": raw[208:212], "rms": f(54), "nlabl": i(55),
    }
    h["labels"] = [raw[224 + 80 * k:304 + 80 * k].rstrip(b" \x00").decode("ascii", "replace")
                   for k in range(min(max(h["nlabl"], 0), 10))]
    if h["mode"] not in _MRC_MODES:
        raise ValueError(f"unsupported MRC mode {h['mode']}")
    if min(h["nx"], h["ny"], h["nz"]) < 1 or h["nsymbt"] < 0:
        raise ValueError("corrupt MRC header (sizes)")
    mx = h["mx"] if h["mx"] > 0 else h["nx"]
    h["voxel_size"] = tuple(c / m if m > 0 else 1.0 for c, m in zip(h["cella"], (mx, h["my"] or 1, h["mz"] or 1)))
    return h


def read_mrc(path, mmap=True):
    """(data, header): data indexed [column, row, section] (NX, NY, NZ) in
    the file's sample type (memory-mapped read-only when mmap)."""
    with open(path, "rb") as fh:
        raw = fh.read(1024)
    h = read_mrc_header(raw)
    dt = np.dtype(_MRC_MODES[h["mode"]]).newbyteorder(h["endian"])
    count = h["nx"] * h["ny"] * h["nz"]
    offset = 1024 + h["nsymbt"]
    if mmap:
        a = np.memmap(path, dtype=dt, mode="r", offset=offset, shape=(count,))
    else:
        a = np.fromfile(path, dtype=dt, count=count, offset=offset)
    if a.size != count:
        raise ValueError("MRC file shorter than its header says")
    return a.reshape(h["nz"], h["ny"], h["nx"]).transpose(2, 1, 0), h


def write_mrc(path, data, voxel_size=1.0, labels=()):
    """Write data indexed [column, row, section] as an MRC2014 file
    (little-endian; float32 -> mode 2, int16 -> 1, int8 -> 0,
    uint16 -> 6, float16 -> 12, complex64 -> 4; float64 is stored as
    float32)."""
    a = np.asarray(data)
    if a.ndim == 2:
        a = a[:, :, None]
    if a.ndim != 3:
        raise ValueError("MRC data must be 2D or 3D")
    if a.dtype == np.float64:
        a = a.astype(np.float32)
    if a.dtype not in _MRC_CODES:
        raise ValueError(f"no MRC mode for {a.dtype}")
    nx, ny, nz = a.shape
    vs = np.broadcast_to(np.asarray(voxel_size, float), (3,))
    real = np.abs(a) if np.iscomplexobj(a) else a
    words = bytearray(1024)
    struct.pack_into("<10i", words, 0, nx, ny, nz, _MRC_CODES[a.dtype], 0, 0, 0, nx, ny, nz)
    struct.pack_into("<6f", words, 40, nx * vs[0], ny * vs[1], nz * vs[2], 90.0, 90.0, 90.0)
    struct.pack_into("<3i", words, 64, 1, 2, 3)
    struct.pack_into("<3f", words, 76, float(real.min()), float(real.max()), float(real.mean()))
    struct.pack_into("<2i", words, 88, 1 if nz > 1 else 0, 0)
    struct.pack_into("<i", words, 108, 20140)
    words[208:212] = b"MAP "
    words[212:216] = b"\x44\x44\x00\x00"
    struct.pack_into("<f", words, 216, float(real.std()))
    labels = list(labels)[:10]
    struct.pack_into("<i", words, 220, len(labels))
    for k, lab in enumerate(labels):
        b = lab.encode("ascii", "replace")[:80]
        words[224 + 80 * k:224 + 80 * k + len(b)] = b
    with open(path, "wb") as fh:
        fh.write(bytes(words))
        fh.write(np.ascontiguousarray(a.transpose(2, 1, 0)).astype(a.dtype.newbyteorder("<"), copy=False).tobytes())


# ---------------------------------------------------------------------------
# Multi-page TIFF
# ---------------------------------------------------------------------------
_TIFF_TYPES = {1: ("B", 1), 2: ("c", 1), 3: ("H", 2), 4: ("I", 4), 6: ("b", 1), 8: ("h", 2), 9: ("i", 4),
               11: ("f", 4), 12: ("d", 8), 16: ("Q", 8)}
_SAMPLE_KIND = {1: "u", 2: "i", 3: "f"}


def _tiff_pages(buf):
    e = {b"II": "<", b"MM": ">"}.get(bytes(buf[:2]))
    if e is None or struct.unpack(e + "H", buf[2:4])[0] != 42:
        raise ValueError("not a classic TIFF file")
    off = struct.unpack(e + "I", buf[4:8])[0]
    seen = set()
    while off:
        if off in seen or off + 2 > len(buf):
            raise ValueError("corrupt TIFF IFD chain")
        seen.add(off)
        n = struct.unpack(e + "H", buf[off:off + 2])[0]
        tags = {}
        for k in range(n):
            ent = off + 2 + 12 * k
            tag, typ, cnt = struct.unpack(e + "HHI", buf[ent:ent + 8])
            if typ not in _TIFF_TYPES:
                continue
            fmt, size = _TIFF_TYPES[typ]
            if fmt == "c":
                continue
            nbytes = size * cnt
            src = ent + 8 if nbytes <= 4 else struct.unpack(e + "I", buf[ent + 8:ent + 12])[0]
            tags[tag] = struct.unpack(e + fmt * cnt, buf[src:src + nbytes])
        yield e, tags
        off = struct.unpack(e + "I", buf[off + 2 + 12 * n:off + 6 + 12 * n])[0]


def read_tiff_stack(path):
    """All pages of a multi-page TIFF as one array [page, row, column]."""
    with open(path, "rb") as fh:
        buf = memoryview(fh.read())
    pages = []
    for e, t in _tiff_pages(buf):
        w, hgt = t[256][0], t[257][0]
        bits = t.get(258, (1,))[0]
        if t.get(259, (1,))[0] != 1:
            raise ValueError("compressed TIFF pages are not supported")
        if t.get(277, (1,))[0] != 1:
            raise ValueError("only one sample per pixel is supported")
        if 322 in t:
            raise ValueError("tiled TIFF pages are not supported")
        kind = _SAMPLE_KIND.get(t.get(339, (1,))[0])
        if kind is None or bits not in (8, 16, 32, 64):
            raise ValueError(f"unsupported TIFF samples ({bits} bits, format {t.get(339, (1,))[0]})")
        dt = np.dtype(f"{e}{kind}{bits // 8}")
        data = b"".join(bytes(buf[o:o + c]) for o, c in zip(t[273], t[279]))
        img = np.frombuffer(data, dtype=dt, count=w * hgt)
        pages.append(img.reshape(hgt, w).astype(dt.newbyteorder("="), copy=False))
    if not pages:
        raise ValueError("TIFF file has no pages")
    if any(p.shape != pages[0].shape for p in pages):
        raise ValueError("TIFF pages differ in size")
    return np.stack(pages)


def read_rawtlt(path):
    """Tilt angles of a `.rawtlt` file (one angle in degrees per line) in
    radians, as loadtxt(path) * pi / 180."""
    return np.loadtxt(path, dtype=np.float64).reshape(-1) * np.pi / 180


def load_tilt_series(tif_path, rawtlt_path=None, dtype=np.float32):
    """(data, angles): data[angle, width, slice] = page[angle].T, the layout
    DogBones_recon_small.py:17-24 and Bipyramid_recon.py:57-62 build; angles
    in radians (None without a .rawtlt file)."""
    pages = read_tiff_stack(tif_path)
    data = np.ascontiguousarray(pages.transpose(0, 2, 1), dtype=dtype)
    angles = read_rawtlt(rawtlt_path) if rawtlt_path is not None else None
    if angles is not None and angles.size != data.shape[0]:
        raise ValueError(f"{angles.size} angles for {data.shape[0]} pages")
    return data, angles


# ---------------------------------------------------------------------------
# savemat outputs
# ---------------------------------------------------------------------------
def save_recon(path, Radon, recon, **extra):
    """savemat(path, {'view': Radon.discretise(recon), 'X': x, 'R': r,
    'I': I}) as Bipyramid_recon.py:140-144 writes; extra entries are added
    as given."""
    from scipy.io import savemat

    def host(v):
        v = v.asarray() if hasattr(v, "asarray") else v
        return v.detach().cpu().numpy() if hasattr(v, "detach") else np.asarray(v)

    out = {"view": host(Radon.discretise(recon)), "X": host(recon.x), "R": host(recon.r), "I": host(recon.I)}
    out.update({k: host(v) for k, v in extra.items()})
    savemat(path, out)
    return out
