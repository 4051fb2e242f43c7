"""CPU tests of the C ABI surface and the host-side logic (no GPU compute)."""
import ctypes
import importlib
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(gt_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(gt):
    lib = gt._lib.load()
    declared = _declared_functions()
    assert len(declared) >= 12
    for name in declared:
        assert hasattr(lib, name), name
        assert name in gt._lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.gt_version().startswith(b"gausstomo")


def test_argument_validation_errors(gt):
    """Bad arguments are refused before any device work, with a message."""
    L = gt._lib
    lib = L.load()
    a = L.gt_atoms(3, 5, 0, 1, 1, 1)
    g = L.gt_geom(3, 3, 4, 8, 8, 1, 1, 1, 1, 1, 1)
    bad_dim = L.gt_atoms(4, 5, 0, 1, 1, 1)
    assert lib.gt_forward(ctypes.byref(bad_dim), ctypes.byref(g), 1, 1, 1 << 20, None) == -1
    assert b"dim" in lib.gt_last_error()
    g2 = L.gt_geom(2, 2, 4, 8, 1, 1, 1, 1, None, 1, None)
    assert lib.gt_forward(ctypes.byref(a), ctypes.byref(g2), 1, 1, 1 << 20, None) == -1
    assert b"does not match" in lib.gt_last_error()
    g3 = L.gt_geom(3, 3, 4, 8, 8, 1, 1, 1, None, 1, 1)  # lenT 3 needs w1
    assert lib.gt_forward(ctypes.byref(a), ctypes.byref(g3), 1, 1, 1 << 20, None) == -1
    assert lib.gt_derivs(ctypes.byref(a), ctypes.byref(g), 1, 7, 2, None, 1, None, 1, 1 << 20, None) == -1
    assert b"mode" in lib.gt_last_error()
    need = lib.gt_forward_workspace(ctypes.byref(a), ctypes.byref(g))
    assert need >= 5 * 4 * 64
    assert lib.gt_forward(ctypes.byref(a), ctypes.byref(g), 1, 256, need - 1, None) == -3
    assert b"workspace" in lib.gt_last_error()
    # the no-allocation step refuses a sweep beyond its scratch array before any device call
    assert lib.gt_noalloc_scratch_bytes() == 128 << 20
    big = L.gt_atoms(3, 20000, 0, 1, 1, 1)
    gbig = L.gt_geom(3, 3, 120, 16, 16, 1, 1, 1, 1, 1, 1)
    assert lib.gt_derivs_workspace(ctypes.byref(big), ctypes.byref(gbig), 2) > 128 << 20
    assert lib.gt_forward_derivs_noalloc(ctypes.byref(big), ctypes.byref(gbig), 1, 2, 1, None, 1, None, None) == -3
    assert b"scratch" in lib.gt_last_error()
    # an unknown point-spread cell rule is refused
    assert lib.gt_psf_forward_rule(ctypes.byref(a), ctypes.byref(g), 5, 1, 1, 1 << 20, None) == -1
    assert b"rule" in lib.gt_last_error()
    assert lib.gt_psf_derivs_rule(ctypes.byref(a), ctypes.byref(g), 1, 1, 2, 9, None, 1, None, 1, 1 << 20, None) == -1
    assert b"rule" in lib.gt_last_error()
    assert lib.gt_psf_volume_rule(ctypes.byref(a), None, -1, 1, 1, 1 << 20, None) == -1
    assert b"rule" in lib.gt_last_error()
    psf = importlib.import_module("2018_dictionary_tomography_amd.psf")
    assert psf.check_rule("reference") == psf.check_rule("sampled") == L.GT_PSF_SAMPLED
    with pytest.raises(ValueError):
        psf.check_rule("midpoint")
    with pytest.raises(L.GTError):
        L.check(-1, "probe")


def test_atom_space_random_matches_reference_draw_order():
    """AtomSpace.random draws r, then x, then I from numpy's legacy generator
    (dictionary_def.py:130-139): r ~ U[1, 2) with the off-diagonal slots
    divided by 4, x ~ U[-0.8, 0.8), I ~ U[0, 2); the seed-3 stream written
    out by hand below."""
    sp = importlib.import_module("2018_dictionary_tomography_amd.spaces")
    torch = importlib.import_module("torch")
    for iso, D in ((False, 3), (True, 3), (False, 2), (True, 2)):
        space = sp.AtomSpace(D, isotropic=iso)
        a = space.random(4, seed=3, device=torch.device("cpu"))
        rs = np.random.RandomState(3)
        nr = 1 if iso else D * (D + 1) // 2
        r = rs.rand(4, nr) + 1
        if not iso:
            r[:, D:] /= 4
        x = 0.8 * (2 * rs.rand(4, D) - 1)
        I = 2 * rs.rand(4)
        assert a.r.shape == (4, nr) and a.x.shape == (4, D) and a.I.shape == (4,)
        assert np.array_equal(a.r.numpy(), np.float32(r)) and np.array_equal(a.x.numpy(), np.float32(x))
        assert np.array_equal(a.I.numpy(), np.float32(I))
    # the first draws of the stream itself (numpy legacy MT19937, seed 3)
    assert np.allclose(np.random.RandomState(3).rand(3), [0.5507979, 0.70814782, 0.29090474])


def test_geometry_helpers():
    sp = importlib.import_module("2018_dictionary_tomography_amd.spaces")
    c = sp.cell_centres(-1, 1, 4)
    assert np.allclose(c, [-0.75, -0.25, 0.25, 0.75])
    assert np.allclose(sp.boundary_nodes(0, 1, 3), [0, 0.5, 1])
    T, (w0, w1) = sp.axis_tilt_geometry(np.linspace(-1, 1, 5), [0, 1, 0], ref=np.array([1.0, 0, 0]))
    for u, v in ((T, w0), (T, w1), (w0, w1)):
        assert np.allclose((u * v).sum(1), 0)
    assert np.allclose(np.cross(w0, w1), T)
    ps = sp.ProjSpace(angles=np.linspace(0, np.pi, 7), detector=(np.linspace(-1, 1, 5), np.linspace(-1, 1, 3)))
    assert ps.shape == [7, 5, 3] and ps.dim == 3 and ps.orientations.shape == (7, 2)
    ps2 = sp.ProjSpace(angles=np.linspace(0, np.pi, 7), detector=np.linspace(-1, 1, 5))
    assert ps2.shape == [7, 5] and ps2.dim == 2


def test_product_geometry_equals_pinned_oracle_geometry(oracle):
    """The product's theta_to_vec / perp (what every GPU test and the drop-in
    ProjSpace feed the kernels) equal the oracle's restatements, which
    tests/test_oracle.py pins (orthonormal frames, the explicit forms of
    atomFuncs.py:13-74), for one angle vector (2D) and two (3D Euler pairs)."""
    sp = importlib.import_module("2018_dictionary_tomography_amd.spaces")
    rng = np.random.default_rng(2)
    for theta in ([rng.uniform(-3, 3, 11)], [rng.uniform(-3, 3, 5), rng.uniform(-1.5, 1.5, 4)]):
        T = sp.theta_to_vec(theta)
        To = oracle.theta_to_vec(theta)
        assert np.allclose(np.asarray(T).reshape(To.shape), To, rtol=0, atol=1e-15)
        W, Wo = sp.perp(theta), oracle.perp(theta)
        assert len(W) == len(Wo)
        for w, wo in zip(W, Wo):
            assert np.allclose(np.asarray(w).reshape(wo.shape), wo, rtol=0, atol=1e-15)


def test_workload_cholesky_convention():
    """U = L^T / sqrt(60): 30 |U y|^2 = y^T L L^T y / 2."""
    W = importlib.import_module("2018_dictionary_tomography_amd.workloads")
    rng = np.random.default_rng(0)
    I, x, r, L = W.draw_aniso3(5, rng)
    for j in range(5):
        U = np.array([[r[j, 0], r[j, 3], r[j, 5]], [0, r[j, 1], r[j, 4]], [0, 0, r[j, 2]]])
        y = rng.normal(size=3)
        assert np.isclose(30 * np.sum((U @ y) ** 2), 0.5 * y @ L[j] @ L[j].T @ y)
    assert np.all((x >= -0.7) & (x <= 0.7)) and np.all((I >= 0.5) & (I <= 1.5))


def test_struct_layouts_match_header(tmp_path):
    """The ctypes mirrors of every struct of include/gausstomo.h have its C
    layout (offset of every field, and the size), checked by compiling the
    header with the host C compiler."""
    import shutil
    import subprocess

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    F = importlib.import_module("2018_dictionary_tomography_amd.fit")
    L = importlib.import_module("2018_dictionary_tomography_amd._lib")
    checks = []
    for cls, cname in ((F.gt_fit_slot, "gt_fit_slot"), (F.gt_fit_block, "gt_fit_block"), (L.gt_atoms, "gt_atoms"),
                       (L.gt_geom, "gt_geom"), (L.gt_vol, "gt_vol"), (L.gt_kgrid, "gt_kgrid"),
                       (L.gt_kplanes, "gt_kplanes"), (L.gt_kplanes64, "gt_kplanes64"), (L.gt_atoms64, "gt_atoms64")):
        checks.append((f"sizeof({cname})", ctypes.sizeof(cls)))
        for name, _ in cls._fields_:
            checks.append((f"offsetof({cname}, {name})", getattr(cls, name).offset))
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"gausstomo.h\"\nint main(void) {\n"
    src += "".join(f'    printf("%zu\\n", (size_t){expr});\n' for expr, _ in checks) + "    return 0;\n}\n"
    (tmp_path / "layout.c").write_text(src)
    exe = tmp_path / "layout"
    subprocess.run([cc, "-I", os.path.join(ROOT, "include"), str(tmp_path / "layout.c"), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    for (expr, want), g in zip(checks, got):
        assert g == want, (expr, g, want)
