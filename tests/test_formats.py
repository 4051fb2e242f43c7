"""On-disk formats: MRC2014 read/write, multi-page TIFF tilt series (against
PIL, the reader the reference's scripts use), .rawtlt angles and the
savemat output of the reconstruction scripts.
"""
import importlib
import struct
import types

import numpy as np
import pytest

fm = importlib.import_module("2018_dictionary_tomography_amd.formats")


def _mrc_bytes(data_xyz, mode, endian, ext=b"", cell=(10.0, 20.0, 30.0)):
    """An MRC2014 file assembled word by word from the published layout:
    words 1-10 sizes/mode/starts/sampling, 11-16 cell, 17-19 axis map,
    20-22 min/max/mean, 23 space group, 24 extended header size, 28 version,
    53 'MAP ', 54 machine stamp, 56 label count, labels from byte 224."""
    nx, ny, nz = data_xyz.shape
    hdr = bytearray(1024)
    ints = {0: nx, 1: ny, 2: nz, 3: mode, 7: nx, 8: ny, 9: nz, 16: 1, 17: 2, 18: 3, 22: 1, 23: len(ext),
            27: 20140, 55: 1}
    floats = {10: cell[0], 11: cell[1], 12: cell[2], 13: 90.0, 14: 90.0, 15: 90.0, 19: -1.0, 20: 1.0, 21: 0.0}
    for k, v in ints.items():
        struct.pack_into(endian + "i", hdr, 4 * k, v)
    for k, v in floats.items():
        struct.pack_into(endian + "f", hdr, 4 * k, v)
    hdr[208:212] = b"MAP "
    hdr[212:214] = b"\x44\x44" if endian == "<" else b"\x11\x11"
    hdr[224:234] = b"test label"
    body = np.ascontiguousarray(data_xyz.transpose(2, 1, 0)).astype(data_xyz.dtype.newbyteorder(endian)).tobytes()
    return bytes(hdr) + ext + body


@pytest.mark.parametrize("endian", ["<", ">"])
@pytest.mark.parametrize("dtype,mode", [(np.float32, 2), (np.int16, 1), (np.int8, 0), (np.uint16, 6),
                                        (np.float16, 12)])
def test_read_mrc_from_spec_bytes(tmp_path, endian, dtype, mode):
    rng = np.random.default_rng(mode)
    a = (rng.normal(0, 50, (5, 4, 3))).astype(dtype)
    p = tmp_path / "v.mrc"
    p.write_bytes(_mrc_bytes(a, mode, endian, ext=b"\x01" * 96))
    d, h = fm.read_mrc(str(p))
    assert d.shape == (5, 4, 3) and np.array_equal(np.asarray(d), a)
    assert h["nsymbt"] == 96 and h["labels"] == ["test label"] and h["endian"] == endian
    assert np.allclose(h["voxel_size"], (2.0, 5.0, 10.0))
    # a column of the file is the first axis: voxel (1, 0, 0) follows (0, 0, 0)
    raw = p.read_bytes()[1024 + 96:]
    first_two = np.frombuffer(raw[:2 * np.dtype(dtype).itemsize], dtype=np.dtype(dtype).newbyteorder(endian))
    assert first_two[0] == a[0, 0, 0] and first_two[1] == a[1, 0, 0]


def test_write_read_mrc_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    for a in (rng.normal(size=(6, 7, 8)).astype(np.float32), rng.integers(-9, 9, (3, 4, 2)).astype(np.int16),
              rng.normal(size=(4, 5)).astype(np.float64)):
        p = tmp_path / "w.mrc"
        fm.write_mrc(str(p), a, voxel_size=(0.5, 0.5, 2.0), labels=["round trip"])
        d, h = fm.read_mrc(str(p), mmap=False)
        ref = a if a.ndim == 3 else a[:, :, None]
        assert d.shape == ref.shape and np.allclose(d, ref.astype(d.dtype), rtol=0, atol=0)
        assert np.allclose(h["voxel_size"][:2], 0.5) and h["labels"] == ["round trip"]
        assert np.isclose(h["dmax"], np.float32(ref.max()))
    with pytest.raises(ValueError):
        fm.write_mrc(str(tmp_path / "bad.mrc"), np.zeros((2, 2, 2), dtype=np.int64))


@pytest.mark.parametrize("dtype,pil_mode", [(np.uint8, "L"), (np.uint16, "I;16"), (np.int32, "I"),
                                            (np.float32, "F")])
def test_tiff_stack_against_pil(tmp_path, dtype, pil_mode):
    Image = pytest.importorskip("PIL.Image")
    rng = np.random.default_rng(7)
    pages = [(rng.uniform(0, 200, (9, 13))).astype(dtype) for _ in range(4)]
    ims = [Image.fromarray(p, mode=pil_mode) if pil_mode != "I;16" else Image.fromarray(p) for p in pages]
    path = tmp_path / "stack.tif"
    ims[0].save(str(path), save_all=True, append_images=ims[1:])
    got = fm.read_tiff_stack(str(path))
    assert got.shape == (4, 9, 13) and np.array_equal(got.astype(np.float64), np.stack(pages).astype(np.float64))
    # the reference's reading: [asarray(page).T for each page] (DogBones_recon_small.py:19-24)
    with Image.open(str(path)) as im:
        ref = np.asarray([np.asarray(im).T for i in range(im.n_frames) if im.seek(i) is None])
    tlt = tmp_path / "a.rawtlt"
    tlt.write_text("\n".join(str(v) for v in (-60.0, -20.0, 20.0, 60.0)) + "\n")
    data, ang = fm.load_tilt_series(str(path), str(tlt))
    assert data.shape == ref.shape and np.array_equal(data, ref.astype(np.float32))
    assert np.allclose(ang, np.array([-60, -20, 20, 60]) * np.pi / 180)


def test_tiff_rejects_what_it_cannot_read(tmp_path):
    p = tmp_path / "x.tif"
    p.write_bytes(b"II*\x00" + b"\x00" * 8)
    with pytest.raises(ValueError):
        fm.read_tiff_stack(str(p))
    p.write_bytes(b"PK\x03\x04")
    with pytest.raises(ValueError):
        fm.read_tiff_stack(str(p))


def test_save_recon_keys(tmp_path):
    from scipy.io import loadmat
    recon = types.SimpleNamespace(x=np.zeros((2, 3)), r=np.ones((2, 6)), I=np.array([1.0, 2.0]))
    Radon = types.SimpleNamespace(discretise=lambda a: np.full((4, 4, 4), 3.0))
    p = tmp_path / "out.mat"
    fm.save_recon(str(p), Radon, recon)
    m = loadmat(str(p))
    assert set(k for k in m if not k.startswith("__")) == {"view", "X", "R", "I"}
    assert m["view"].shape == (4, 4, 4) and np.array_equal(m["I"].reshape(-1), [1.0, 2.0])
