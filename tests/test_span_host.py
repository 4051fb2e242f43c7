"""Culling predicates of the pixel kernels, checked by brute force on the host.

The derivative kernels walk, per detector column of a footprint box, only the
rows of a span derived from the roots of the column's quadratic (gt_pixel.h,
column_span).  A pixel left out of a span would change the moments by at most
exp(-30) of the atom's peak, far below what the FP64 parity tests resolve, so
this test checks the spans directly: tests/native/span_check.cpp builds
footprint records with the library's own pair_setup for random 3D atoms on a
61-angle tilt series and asserts that every box pixel whose FP32 exponent
passes the cutoff lies in its column's span (uniform detectors; windows of a
few rows; a non-uniform ascending detector, where the whole window is walked),
and that the forward kernel's per-wave test (support_meets_rect) keeps every
8 x 16 rectangle that holds such a pixel.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def span_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("span") / "span_check")
    src = os.path.join(ROOT, "tests", "native", "span_check.cpp")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "2018_dictionary_tomography_amd", "csrc")]
    res = subprocess.run([HIPCC, "-O2", "-std=c++17", *inc, src, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


@pytest.mark.parametrize("atoms,side,window,nonuniform", [(200, 128, 512, 0), (120, 256, 512, 0), (200, 64, 512, 0),
                                                          (60, 256, 16, 0), (100, 128, 512, 1)])
def test_every_support_pixel_is_in_its_column_span(span_check, atoms, side, window, nonuniform):
    res = subprocess.run([span_check, str(atoms), str(side), str(window), str(nonuniform)], capture_output=True,
                         text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    pairs, inside, walked, box, rects, kept = map(int, res.stdout.split())
    assert pairs > 0 and inside > 0
    assert inside <= walked <= box
    assert kept < rects  # the exact test drops rectangles the boxes touch
    if nonuniform:
        assert walked == box  # no grid: every row of the window
    else:
        assert walked < 0.9 * box  # the spans do cut the corners of the boxes


@pytest.fixture(scope="module")
def ft_patch_check(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("ftp") / "ft_patch_check")
    src = os.path.join(ROOT, "tests", "native", "ft_patch_check.cpp")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "2018_dictionary_tomography_amd", "csrc")]
    res = subprocess.run([HIPCC, "-O2", "-std=c++17", *inc, src, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe


@pytest.mark.parametrize("iso", [0, 1])
def test_fourier_patch_test_keeps_every_support_sample(ft_patch_check, iso):
    """The Fourier grid kernel's per-patch test (gt_ftpatch.h, ft_patch_min)
    keeps every 8 x 8 (k1, k2) patch at every k0 of [-64, 64)^3 that holds a
    sample inside an atom's support, and drops most of the others."""
    res = subprocess.run([ft_patch_check, "12", str(iso)], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    patches, kept, with_support = map(int, res.stdout.split())
    assert 0 < with_support <= kept < patches
    assert kept <= 1.05 * with_support + 16


def test_index_searches_match_bisection(tmp_path):
    """The footprint boxes' index searches (gt_pair.h, lower_idx / upper_idx:
    a guess from the end points, checked against its neighbours, bisection
    when the check fails) return what std::lower_bound / std::upper_bound do
    on uniform, non-uniform, clustered and repeated ascending vectors of 1 to
    1000 entries, for queries on, between and beyond the points, +-inf and
    NaN."""
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "idx_check")
    src = os.path.join(ROOT, "tests", "native", "idx_check.cpp")
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(ROOT, "2018_dictionary_tomography_amd", "csrc")]
    res = subprocess.run([HIPCC, "-O2", "-std=c++17", *inc, src, "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert int(res.stdout.split()[1]) > 50000
