"""FP64 oracle for the Gaussian-atom tomography hot path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, and only as the checker.

The heavy loops live in ``gt_oracle.c`` (built into ``libgt_oracle.so`` by
``build()`` below); this module restates the small host-side pieces of the
reference in NumPy:

* ``theta_to_vec`` / ``perp``            GaussDictCode/atomFuncs.py:13-74
* isotropic -> anisotropic slots and the chain-rule collapse of derivatives
  (the intended mathematics of GaussDictCode/bin/NewtonBases.py:530-672,
  873-951, 1072-1152 and GaussDictCode/bin/NewtonBases2D.py:316-326,
  404-441, whose iso paths are broken in the reference)
* ``analytic_gauss_fourier`` / ``analytic_gauss_ifourier``
                                         Fourier_Transform.py:293-376
* ``l2_squared_loss``                    GaussDictCode/transport_loss.py:567-585
* ``newton_linesearch``                  NewtonGaussian.py:18-109 with
                                         GD_lib.py:492-536 (``_step``)
* ``newton_fit`` (also block variant)    NewtonGaussian.py:18-237, any regulariser
* ``gd_linesearch_block``                GD_lib.py:82-178
* ``psf_radproj`` / ``psf_derivs`` / ``psf_volproj``  pixel and voxel
  integrals of the point projector (point_spread_functions.py:168-1350);
  ``psf_radproj_sampled`` / ``psf_derivs_sampled`` / ``psf_volproj_sampled``
  the reference's own sampled rule (:25-88, 201-291, 389-527, 924-1075)
* ``Reg`` (regulariser terms)            GaussDictCode/regularisation.py:22-442;
  values restated from the ``__call__`` methods, gradients and Hessians by
  FP64 automatic differentiation of those values (torch.autograd on the CPU)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgt_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile gt_oracle.c with gcc (OpenMP) into libgt_oracle.so."""
    src = os.path.join(_HERE, "gt_oracle.c")
    if force or not os.path.exists(_LIB_PATH) or os.path.getmtime(_LIB_PATH) < os.path.getmtime(src):
        cmd = ["gcc", "-O2", "-std=c11", "-fopenmp", "-fPIC", "-shared", "-o", _LIB_PATH + ".tmp", src, "-lm"]
        subprocess.check_call(cmd)
        os.replace(_LIB_PATH + ".tmp", _LIB_PATH)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ci = ctypes.c_int
        _lib.or_radproj.argtypes = [ci, ci, dp, dp, dp, ci, ci, dp, dp, dp, ci, dp, ci, dp, dp]
        _lib.or_derivs.argtypes = [ci, ci, ci, ci, dp, dp, dp, ci, ci, dp, dp, dp, ci, dp, ci, dp, dp, dp, dp, dp]
        _lib.or_volproj.argtypes = [ci, ci, dp, dp, dp, ci, dp, ci, dp, ci, dp, dp]
        _lib.or_ft_eval.argtypes = [ci, ci, ci, dp, dp, dp, ctypes.c_long, dp, dp]
        _lib.or_num_threads.restype = ci
    return _lib


def num_threads() -> int:
    return int(lib().or_num_threads())


class CpuBaseline:
    """The dense reference-semantics loops of gt_oracle.c compiled for the
    host it runs on (gcc -O3 -march=native -fopenmp), in FP64 or in FP32 (the
    reference kernels' float32, with single-precision constants), into
    `dest_dir`.  Timed by the benchmark as the CPU baseline; the checker
    build above stays FP64 at -O2."""

    def __init__(self, dest_dir, bits=64):
        src = os.path.join(_HERE, "gt_oracle.c")
        self.bits = int(bits)
        self.path = os.path.join(dest_dir, f"libgt_cpu{self.bits}.so")
        flags = ["-O3", "-march=native", "-std=c11", "-fopenmp", "-fPIC", "-shared"]
        if self.bits == 32:
            flags += ["-DGT_REAL=float", "-fsingle-precision-constant"]
        res = subprocess.run(["gcc", *flags, "-o", self.path, src, "-lm"], capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(res.stderr)
        self.flags = " ".join(flags)
        self.lib = ctypes.CDLL(self.path)
        self.dt = np.float32 if self.bits == 32 else np.float64
        rp = ctypes.POINTER(ctypes.c_float if self.bits == 32 else ctypes.c_double)
        ci = ctypes.c_int
        self.lib.or_radproj.argtypes = [ci, ci, rp, rp, rp, ci, ci, rp, rp, rp, ci, rp, ci, rp, rp]
        self.lib.or_derivs.argtypes = [ci, ci, ci, ci, rp, rp, rp, ci, ci, rp, rp, rp, ci, rp, ci, rp, rp, rp, rp, rp]
        self.lib.or_num_threads.restype = ci
        self._rp = rp

    def threads(self):
        return int(self.lib.or_num_threads())

    def _a(self, a):
        a = np.ascontiguousarray(a, dtype=self.dt)
        return a, a.ctypes.data_as(self._rp)

    def _geom(self, g):
        T, w0 = self._a(g.orientations), self._a(g.ortho[0])
        w1 = self._a(g.ortho[1]) if len(g.ortho) > 1 else (None, None)
        p0 = self._a(g.detector[0])
        p1 = self._a(g.detector[1]) if len(g.detector) > 1 else self._a(np.zeros(1))
        return T, w0, w1, p0, p1

    def radproj(self, I, x, r, g):
        D = g.dim
        (I_, pI), (x_, px), (r_, pr) = self._a(I), self._a(x), self._a(r)
        T, w0, w1, p0, p1 = self._geom(g)
        nA, lenT = T[0].shape
        n0, n1 = p0[0].size, (p1[0].size if D == 3 else 1)
        R, pR = self._a(np.zeros((nA, n0, n1)))
        self.lib.or_radproj(D, I_.size, pI, px, pr, nA, lenT, T[1], w0[1], w1[1], n0, p0[1], n1, p1[1], pR)
        return R

    def derivs(self, I, x, r, g, C, mode="L2", order=2):
        D = g.dim
        (I_, pI), (x_, px), (r_, pr) = self._a(I), self._a(x), self._a(r)
        T, w0, w1, p0, p1 = self._geom(g)
        nA, lenT = T[0].shape
        n0, n1 = p0[0].size, (p1[0].size if D == 3 else 1)
        N, P = I_.size, 1 + D + r_.size // max(I_.size, 1)
        C_, pC = self._a(C)
        (f, pf), (G, pG), (H, pH) = self._a(np.zeros(N)), self._a(np.zeros((N, P))), self._a(np.zeros((N, P, P)))
        self.lib.or_derivs(D, 1 if mode == "L2" else 0, order, N, pI, px, pr, nA, lenT, T[1], w0[1], w1[1], n0, p0[1],
                           n1, p1[1], pC, pf, pG, pH)
        return f, G, H


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---------------------------------------------------------------------------
# Geometry: GaussDictCode/atomFuncs.py:13-74
# ---------------------------------------------------------------------------
def theta_to_vec(theta):
    """Ray directions for a tuple of angle vectors (atomFuncs.py:13-47)."""
    theta = [np.asarray(t, dtype=np.float64) for t in theta]
    if len(theta) == 1:
        t = theta[0]
        return np.stack([-np.sin(t), np.cos(t)], axis=-1)
    if len(theta) == 2:
        t0, t1 = np.meshgrid(theta[0], theta[1], indexing="ij")
        return np.stack([-np.sin(t0) * np.cos(t1), np.cos(t0) * np.cos(t1), np.sin(t1)], axis=-1)
    raise ValueError("only 2D and 3D geometries are restated")


def perp(theta):
    """Detector bases for a tuple of angle vectors (atomFuncs.py:50-74)."""
    theta = [np.asarray(t, dtype=np.float64) for t in theta]
    if len(theta) == 1:
        t = theta[0]
        return (np.stack([np.cos(t), np.sin(t)], axis=-1),)
    if len(theta) == 2:
        t0, t1 = np.meshgrid(theta[0], theta[1], indexing="ij")
        w0 = np.stack([np.cos(t0), np.sin(t0), np.zeros_like(t0)], axis=-1)
        w1 = np.stack([np.sin(t0) * np.sin(t1), -np.cos(t0) * np.sin(t1), np.cos(t1)], axis=-1)
        return w0, w1
    raise ValueError("only 2D and 3D geometries are restated")


# ---------------------------------------------------------------------------
# Atom slot conventions
# ---------------------------------------------------------------------------
def n_slots(D):
    return D * (D + 1) // 2


def to_aniso(r, D):
    """Isotropic radius -> anisotropic slots U = rho * Id (NewtonBases.py:282-291,
    NewtonBases2D.py:404-412 intent)."""
    r = np.asarray(r, dtype=np.float64).reshape(len(r), -1)
    if r.shape[1] == n_slots(D):
        return r
    out = np.zeros((r.shape[0], n_slots(D)))
    out[:, :D] = r[:, :1]
    return out


def iso_collapse(D):
    """d theta_aniso / d theta_iso for theta = [I, x, r]."""
    Pa, Pi = 1 + D + n_slots(D), 1 + D + 1
    E = np.zeros((Pa, Pi))
    for i in range(1 + D):
        E[i, i] = 1.0
    E[1 + D:1 + D + D, 1 + D] = 1.0
    return E


class Geometry:
    """Explicit projection geometry mirroring ProjSpace(orientations=, ortho=, detector=)
    (GaussDictCode/dictionary_def.py:263-283)."""

    def __init__(self, orientations, ortho, detector):
        self.orientations = np.asarray(orientations, dtype=np.float64)
        self.ortho = tuple(np.asarray(o, dtype=np.float64) for o in ortho)
        self.detector = [np.asarray(p, dtype=np.float64).reshape(-1) for p in detector]
        self.shape = [self.orientations.shape[0]] + [p.size for p in self.detector]

    @property
    def dim(self):
        return 2 if len(self.detector) == 1 else 3


def _geom_args(g: Geometry):
    D = g.dim
    T = _c64(g.orientations)
    lenT = T.shape[1]
    w0 = _c64(g.ortho[0])
    w1 = _c64(g.ortho[1]) if (D == 3 and lenT == 3) else None
    p0 = _c64(g.detector[0])
    p1 = _c64(g.detector[1]) if D == 3 else np.zeros(1)
    return D, T, lenT, w0, w1, p0, p1


def radproj(I, x, r, geom: Geometry):
    """Forward projection in FP64 (NewtonBases.py:714-756, NewtonBases2D.py:274-296)."""
    D, T, lenT, w0, w1, p0, p1 = _geom_args(geom)
    I, x = _c64(I).reshape(-1), _c64(x).reshape(-1, D)
    ra = _c64(to_aniso(r, D))
    N = I.size
    R = np.zeros(geom.shape if D == 3 else geom.shape + [1])
    lib().or_radproj(D, N, _p(I), _p(x), _p(ra), T.shape[0], lenT, _p(T), _p(w0),
                     _p(w1) if w1 is not None else None, p0.size, _p(p0), p1.size, _p(p1), _p(R))
    return R.reshape(geom.shape)


def derivs(I, x, r, geom: Geometry, C, mode="L2", order=2):
    """Per-atom (f, Df, DDf) over all rays, each atom independent.

    mode "L2":     L2derivs_RadProj (NewtonBases.py:759-870, NewtonBases2D.py:299-401)
    mode "linear": derivs_RadProj   (NewtonBases.py:954-1069, NewtonBases2D.py:415-515)
    Isotropic atoms are differentiated as U = rho * Id and collapsed by the chain rule.
    """
    D, T, lenT, w0, w1, p0, p1 = _geom_args(geom)
    I, x = _c64(I).reshape(-1), _c64(x).reshape(-1, D)
    r = np.asarray(r, dtype=np.float64).reshape(I.size, -1)
    iso = r.shape[1] == 1
    ra = _c64(to_aniso(r, D))
    N = I.size
    P = 1 + D + n_slots(D)
    C = _c64(C)
    f = np.zeros(N)
    Df = np.zeros((N, P))
    DDf = np.zeros((N, P, P))
    lib().or_derivs(D, 1 if mode == "L2" else 0, order, N, _p(I), _p(x), _p(ra), T.shape[0], lenT,
                    _p(T), _p(w0), _p(w1) if w1 is not None else None, p0.size, _p(p0), p1.size,
                    _p(p1), _p(C), _p(f), _p(Df), _p(DDf))
    if iso:
        E = iso_collapse(D)
        Df = Df @ E
        DDf = np.einsum("ai,nab,bj->nij", E, DDf, E)
    return f, Df, DDf


def volproj(I, x, r, coords):
    """Volume render (NewtonBases.py:691-711, NewtonBases2D.py:252-271)."""
    D = len(coords)
    I, x = _c64(I).reshape(-1), _c64(x).reshape(-1, D)
    ra = _c64(to_aniso(r, D))
    y = [_c64(c).reshape(-1) for c in coords]
    if D == 2:
        y.append(np.zeros(1))
    u = np.zeros(tuple(len(c) for c in y))
    lib().or_volproj(D, I.size, _p(I), _p(x), _p(ra), y[0].size, _p(y[0]), y[1].size, _p(y[1]),
                     y[2].size, _p(y[2]), _p(u))
    return u.reshape(tuple(len(c) for c in coords))


def ft_points(I, x, r, K, iso=False):
    """Fourier evaluation at explicit points K (npts x D) (FTGaussian3D.py:95-196,
    FTGaussian2D.py:42-100).  I, x, r are FT-space parameters."""
    K = _c64(K)
    D = K.shape[1]
    I, x = _c64(I).reshape(-1), _c64(x).reshape(-1, D)
    r = _c64(r).reshape(I.size, -1)
    out = np.zeros((K.shape[0], 2))
    lib().or_ft_eval(D, 1 if iso else 0, I.size, _p(I), _p(x), _p(r), K.shape[0], _p(K), _p(out))
    return out[:, 0] + 1j * out[:, 1]


def ft_grid(I, x, r, k, iso=False):
    """Evaluation on a separable grid k = (k0, k1[, k2]) (FTGaussian3D.py:40-50)."""
    mesh = np.meshgrid(*[np.asarray(v, dtype=np.float64) for v in k], indexing="ij")
    K = np.stack([m.reshape(-1) for m in mesh], axis=1)
    return ft_points(I, x, r, K, iso).reshape(mesh[0].shape)


def ft_hyperplane(I, x, r, k0, k1, w0, w1=None, iso=False):
    """Fourier-slice samples k = k0[i1] w0[i0] + k1[i2] w1[i0] (FTGaussian3D.py:113-148);
    w1=None selects the 2.5D form k = (k0 w0[i0,0], k0 w0[i0,1], k1) (:132-148)."""
    k0, k1 = np.asarray(k0, float), np.asarray(k1, float)
    w0 = np.asarray(w0, float)
    nA = w0.shape[0]
    a, i, j = np.meshgrid(np.arange(nA), np.arange(k0.size), np.arange(k1.size), indexing="ij")
    if w1 is None:
        K = np.stack([k0[i] * w0[a, 0], k0[i] * w0[a, 1], k1[j]], axis=-1)
    else:
        w1 = np.asarray(w1, float)
        K = k0[i][..., None] * w0[a] + k1[j][..., None] * w1[a]
    return ft_points(I, x, r, K.reshape(-1, 3), iso).reshape(nA, k0.size, k1.size)


# ---------------------------------------------------------------------------
# Fourier parameter maps: Fourier_Transform.py:293-376
# ---------------------------------------------------------------------------
def analytic_gauss_fourier(I, x, r):
    """(I, x, r) of f(y) = I exp(-|U(y-x)|^2/2) -> FT parameters (Fourier_Transform.py:293-333)."""
    I, x, r = np.asarray(I, float), np.asarray(x, float), np.asarray(r, float)
    N, D = x.shape
    tau = 2 * np.pi
    nI, nx, nr = np.empty(N), tau * x, np.empty_like(r)
    if r.shape[1] == 1:
        nI[:] = I * np.sqrt(tau) ** D * r[:, 0] ** D
        nr[:, 0] = tau / r[:, 0]
        return nI, nx, nr
    det = np.prod(r[:, :D], axis=1)
    nI[:] = I * np.sqrt(tau) ** D / det
    if D == 2:
        nr[:, 0] = tau / r[:, 0]
        nr[:, 1] = tau / r[:, 1]
        nr[:, 2] = -(tau * r[:, 2]) / (r[:, 0] * r[:, 1])
    else:
        nr[:, 0] = tau / r[:, 0]
        nr[:, 1] = tau / r[:, 1]
        nr[:, 2] = tau / r[:, 2]
        nr[:, 3] = -(tau * r[:, 3]) / (r[:, 0] * r[:, 1])
        nr[:, 4] = -(tau * r[:, 4]) / (r[:, 1] * r[:, 2])
        nr[:, 5] = tau * (r[:, 3] * r[:, 4] / r[:, 1] - r[:, 5]) / (r[:, 0] * r[:, 2])
    return nI, nx, nr


def analytic_gauss_ifourier(I, x, r):
    """Inverse parameter map (Fourier_Transform.py:336-376)."""
    I, x, r = np.asarray(I, float), np.asarray(x, float), np.asarray(r, float)
    N, D = x.shape
    tau = 2 * np.pi
    nx = x / tau
    nr = np.empty_like(r)
    if r.shape[1] == 1:
        nr[:, 0] = tau / r[:, 0]
        nI = I / (nr[:, 0] ** D * np.sqrt(tau) ** D)
        return nI, nx, nr
    if D == 2:
        nr[:, 0] = tau / r[:, 0]
        nr[:, 1] = tau / r[:, 1]
        nr[:, 2] = -(tau * r[:, 2]) / (r[:, 0] * r[:, 1])
    else:
        nr[:, 0] = tau / r[:, 0]
        nr[:, 1] = tau / r[:, 1]
        nr[:, 2] = tau / r[:, 2]
        nr[:, 3] = -(tau * r[:, 3]) / (r[:, 0] * r[:, 1])
        nr[:, 4] = -(tau * r[:, 4]) / (r[:, 1] * r[:, 2])
        nr[:, 5] = tau * (r[:, 3] * r[:, 4] / r[:, 1] - r[:, 5]) / (r[:, 0] * r[:, 2])
    nI = I * np.prod(nr[:, :D], axis=1) / np.sqrt(tau) ** D
    return nI, nx, nr


# ---------------------------------------------------------------------------
# Fitters
# ---------------------------------------------------------------------------
def l2_squared_loss(R, data):
    """1/2 ||R - data||^2 (transport_loss.py:567-585), here in FP64 (complex
    data: the squared modulus)."""
    R, data = np.asarray(R), np.asarray(data)
    if np.iscomplexobj(R) or np.iscomplexobj(data):
        d = R.astype(complex) - data.astype(complex)
        return 0.5 * float(np.sum(d.real ** 2 + d.imag ** 2))
    d = R.astype(float) - data.astype(float)
    return 0.5 * float(np.sum(d * d))


def _project(D, iso, I, x, r, d):
    """Apply one update with the projections of GD_lib._step (GD_lib.py:502-522)."""
    I2 = max(0.0, I + d[0])
    x2 = x + d[1:1 + D]
    if iso:
        r2 = np.abs(r + d[1 + D])
    else:
        r2 = r + d[1 + D:]
        if D == 2:
            if r2[0] < 0:
                r2[0], r2[2] = -r2[0], -r2[2]
            if r2[1] < 0:
                r2[1] = -r2[1]
        else:
            if r2[0] < 0:
                r2[0], r2[3], r2[5] = -r2[0], -r2[3], -r2[5]
            if r2[1] < 0:
                r2[1], r2[4] = -r2[1], -r2[4]
            if r2[2] < 0:
                r2[2] = -r2[2]
    return I2, x2, r2


def newton_linesearch(I, x, r, geom, data, max_iter, tol=1e-4, min_iter=1, seed=1):
    """Per-atom damped Newton with accept/reject (NewtonGaussian.py:18-109 and
    GD_lib.py:492-536), null regulariser, no initial guess.  FP64 throughout.

    Returns (I, x, r, E) with E the energy history."""
    D = geom.dim
    I, x = np.array(I, float).reshape(-1), np.array(x, float).reshape(-1, D)
    r = np.array(r, float).reshape(I.size, -1)
    iso = r.shape[1] == 1
    n = I.size
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    eps = 1e-4
    h = [1.0] * n
    data = np.asarray(data, float)
    R = radproj(I, x, r, geom)
    E = [l2_squared_loss(R, data)]
    rs = np.random.RandomState(seed)
    for it in range(max_iter[0]):
        brk = [True, True]
        for j in rs.permutation(np.arange(n)):
            for _ in range(max_iter[1]):
                brk[1] = True
                Rj = radproj(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom)
                C = data + Rj - R
                _, df, ddf = derivs(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom, C, "L2", 2)
                df, ddf = df[0].copy(), ddf[0].copy()
                ddf[np.diag_indices_from(ddf)] += 1.0 / h[j]
                dx = np.linalg.solve(ddf, -df)
                # _step
                old = (I[j], x[j].copy(), r[j].copy())
                I[j], x[j], r[j] = _project(D, iso, I[j], x[j], r[j], dx)
                Rnew = R + radproj(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom) - Rj
                En = l2_squared_loss(Rnew, data)
                if En > E[-1]:
                    I[j], x[j], r[j] = old
                    E.append(E[-1])
                    if h[j] > 2 * eps:
                        brk = [False, False]
                    h[j] /= 10
                else:
                    R = Rnew
                    E.append(En)
                    if np.sqrt(np.sum(dx * dx)) > tol:
                        brk = [False, False]
                    h[j] *= 2
                h[j] = min(max(h[j], eps), 1 / eps)
                if brk[1]:
                    break
        if brk[0] and it > min_iter:
            break
    return I, x, r, np.array(E)


def _projector(kind, n_sub):
    """(radproj, derivs) of the point projector (GaussTomo), of the
    pixel-integrated one (SingleParticleGaussTomo, n_sub nodes per pixel
    and axis), of the reference's sampled point-spread rule
    ('psf_reference'), or of the Fourier-slice planes (GaussFTVolume: 'geometry' is
    the sample tuple (k0, k1, w0, w1), values cut at |L k|^2 < 30 and L2
    derivatives at 60, FT-space atoms).  A pair of callables (radproj(I, x,
    r, geom), derivs(I, x, r, geom, C, mode, order)) is used as given: the
    fitters call nothing else, as the reference's fitters only call
    Radon(.) and Radon.L2_derivs."""
    if isinstance(kind, tuple):
        return kind
    if kind == "point":
        return radproj, derivs
    if kind == "psf":
        return (lambda I, x, r, g: psf_radproj(I, x, r, g, n_sub),
                lambda I, x, r, g, C, mode, order: psf_derivs(I, x, r, g, C, mode, order, n_sub))
    if kind == "psf_reference":
        return psf_radproj_sampled, psf_derivs_sampled
    if kind == "ft":
        def fwd(I, x, r, g):
            return ft_hyperplane(I, x, r, g[0], g[1], g[2], g[3], iso=np.asarray(r).shape[-1] == 1)

        def der(I, x, r, g, C, mode, order):
            if mode != "L2":
                raise ValueError("Fourier-plane derivatives are L2 only")
            return ft_derivs(I, x, r, g[0], g[1], g[2], g[3], C, iso=np.asarray(r).shape[-1] == 1)
        return fwd, der
    raise ValueError(kind)


def _atom_dim(projector, geom):
    """Spatial dimension of the atoms: 3 for Fourier-space atoms, else the
    geometry's (a custom projector's geometry only needs a `dim`)."""
    return 3 if isinstance(projector, str) and projector == "ft" else geom.dim


def gd_linesearch_block(I, x, r, geom, data, max_iter, dim="xrI", tol=1e-4, min_iter=1, seed=1, projector="point",
                        n_sub=12, guess=None, return_F=False, trace=None):
    """Block gradient descent (GD_lib.py:82-178), null regulariser, FP64.

    Per atom j of each outer iteration's permutation and per block t of dim
    (filtered to the order x, r, I, :84-92): gradient of the block, step size
    h[t][j] = min(h[t][j], 1/max(|grad|, eps)) with eps = 1e-8 and h starting
    at 1e-2 (:96-99, :149-150), T -= h grad, full energy of the proposal;
    reject (E rises): revert, h /= 2; accept: h *= 1.1 and the sweep is not
    converged if |old - T| > tol |old|; then h = max(h, eps) (:157-166).  No
    projections.  Exit after an outer iteration with no move once its index
    exceeds min_iter (:171).

    Divergences from the reference, each a repair:
    * block t steps along that block's own gradient: I along dF/dI, x along
      dF/dx, r along dF/dr.  The reference picks the gradient by the block's
      position (grad = df[0], df[1:1+D], df[1+D:] for t = 0, 1, 2) while its
      blocks are ordered x, r, I, and then moves every entry of the block by
      the scalar h*grad[0] (:138-151): for dim='xrI' x moves along
      (1, .., 1) dF/dI, r along (1, .., 1) dF/dx0, I by dF/dr0;
    * on reject the sinogram is recomputed from the reverted atoms
      (:157-160 leave the rejected sinogram in place).
    guess: the greedy insertion path as in newton_fit (GD_lib.py:113-126).
    Returns (I, x, r, E) or, with return_F, also F (the fidelity of every
    proposal).  trace (a list): one dict per block step appended, with the
    atom j, block name, step size after the update, accept flag and the two
    BREAK flags after it."""
    D = _atom_dim(projector, geom)
    I, x = np.array(I, float).reshape(-1), np.array(x, float).reshape(-1, D)
    r = np.array(r, float).reshape(I.size, -1)
    n_all = I.size
    dim = "".join(c for c in "xrI" if c.lower() in dim.lower())
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    max_iter = list(max_iter) + [1] * (2 - len(max_iter))
    eps = 1e-8
    h = [[1e-2] * n_all for _ in dim]
    data = np.asarray(data)
    data = data.astype(complex) if np.iscomplexobj(data) else data.astype(float)
    radproj, derivs = _projector(projector, n_sub)
    rs = np.random.RandomState(seed)
    arrs = {"I": I, "x": x, "r": r}
    E, F = [], []
    if guess is None:
        n = n_all
        R = radproj(I, x, r, geom)
        E = [l2_squared_loss(R, data)]
        F = [E[0]]
    else:
        n = 1
    while n <= n_all + 1:
        if n == n_all + 1:
            if guess is None or len(max_iter) == 2:
                break
            n = n_all
            max_iter = [max_iter[2], max_iter[1]]
        elif guess is not None:
            res = data - radproj(I[:n], x[:n], r[:n], geom)
            Ig, xg, rg = guess(res, I[n - 1:n].copy(), x[n - 1:n].copy(), r[n - 1:n].copy())
            I[n - 1], x[n - 1], r[n - 1] = np.reshape(Ig, -1)[0], np.reshape(xg, -1), np.reshape(rg, -1)
            R = radproj(I[:n], x[:n], r[:n], geom)
            Ev = l2_squared_loss(R, data)
            if E:
                E[-1], F[-1] = Ev, Ev
            else:
                E, F = [Ev], [Ev]
        for it in range(max_iter[0]):
            brk = [True, True]
            for j in rs.permutation(np.arange(n)):
                for _ in range(max_iter[1]):
                    brk[1] = True
                    for t, name in enumerate(dim):
                        Rj = radproj(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom)
                        _, df, _ = derivs(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom, data + Rj - R, "L2", 1)
                        df = df[0]
                        if name == "I":
                            grad = df[0:1]
                        elif name == "x":
                            grad = df[1:1 + D]
                        else:
                            grad = df[1 + D:]
                        A = arrs[name]
                        old = np.array(A[j], dtype=float, copy=True)
                        h[t][j] = min(h[t][j], 1.0 / max(np.sqrt(np.sum(grad * grad)), eps))
                        new = old - h[t][j] * grad if name != "I" else old - h[t][j] * grad[0]
                        A[j] = new
                        Rn = radproj(I[:n], x[:n], r[:n], geom)
                        En = l2_squared_loss(Rn, data)
                        F.append(En)
                        acc = not En > E[-1]
                        if not acc:
                            A[j] = old
                            E.append(E[-1])
                            h[t][j] /= 2
                        else:
                            R = Rn
                            E.append(En)
                            h[t][j] *= 1.1
                            if np.sqrt(np.sum((old - new) ** 2)) > tol * np.sqrt(np.sum(old * old)):
                                brk = [False, False]
                        h[t][j] = max(h[t][j], eps)
                        if trace is not None:
                            trace.append({"j": int(j), "block": name, "h": h[t][j], "accept": acc,
                                          "brk": tuple(brk)})
                    if brk[1]:
                        break
            if brk[0] and it > min_iter:
                break
        n += 1
    if return_F:
        return I, x, r, np.array(E), np.array(F)
    return I, x, r, np.array(E)


# ---------------------------------------------------------------------------
# Regularisers, GaussDictCode/regularisation.py:22-442
# ---------------------------------------------------------------------------
class Reg:
    """One regulariser, built from the reference's constructor arguments:

    Reg("null", dim) | Reg("Joubert", dim, intensity, location, (r0, r1))
    | Reg("Mass", dim, m, k=2, s=1) | Reg("Volume", dim, r0, k=2, s=1)
    | Reg("Radius", dim, r0, s, k) | Reg("Iso", dim, s)

    term(th) is the value of one atom's term (th = [I, x, r], one r if iso) as
    the reference's __call__ sums it; derivs(th) its gradient and Hessian by
    automatic differentiation.  Domain tests as in the reference, except that
    Joubert's looks at the diagonal slots only (its logarithm's argument;
    regularisation.py:63-64 tests every slot)."""

    def __init__(self, name, dim, *args):
        self.name, self.dim, self.args = name, dim, args
        if name == "Joubert":
            intensity, location, radius = args
            self.p = (intensity, location, radius[1], radius[1] * radius[0],
                      float(np.prod(radius)) * (1 - np.log(radius[0])))          # :41-51
        elif name in ("Mass", "Volume"):
            a = list(args) + [2, 1][len(args) - 1:]
            self.p = (abs(a[0]) ** dim, a[1], a[2]) if name == "Volume" else tuple(a)  # :111-117, :213-217
        elif name == "Radius":
            r0, s_, k = args
            self.p = (r0, s_, -k)                                                # :321-327
        elif name == "Iso":
            self.p = (args[0],)
        elif name != "null":
            raise ValueError(name)

    def _value(self, th, iso):
        import torch
        D = self.dim
        I, x, r = th[0], th[1:1 + D], th[1 + D:]
        if self.name == "null":
            return th.sum() * 0
        if self.name == "Joubert":                                                # :53-70
            a, b, c, d, e = self.p
            v = r[0] ** D if iso else torch.prod(r[:D])
            return -a * torch.log(I) + (b / 2) * (x * x).sum() + c * v - d * torch.log(v) - e
        if self.name == "Mass":                                                   # :114-129
            m, k, s_ = self.p
            V = I * r[0] ** D if iso else I / torch.prod(r[:D])
            # |V| with the derivative of V at V = 0 (the one-sided derivative the
            # reference's grad formula gives, regularisation.py:131-160)
            return s_ * (torch.where(V >= 0, V, -V) - m) ** k
        if self.name == "Volume":                                                 # :218-236
            p0, k, s_ = self.p
            dd = r[0] ** -D - p0 if iso else torch.prod(r[:D]) - p0
            return s_ * torch.abs(dd) ** k
        if self.name == "Radius":                                                 # :329-343
            r0, s_, kk = self.p
            return s_ * (D * (r[0] - r0) ** kk if iso else ((r[:D] - r0) ** kk).sum())
        if self.name == "Iso":                                                    # :398-407
            return th.sum() * 0 if iso else self.p[0] * (r[D:] * r[D:]).sum() / 2
        raise ValueError(self.name)

    def _outside(self, th, iso):
        D = self.dim
        I, r = th[0], np.asarray(th[1 + D:])
        diag = np.repeat(r[:1], D) if iso else r[:D]
        if self.name == "Joubert":
            return I <= 0 or diag.min() <= 0
        if self.name == "Mass":
            return I < 0
        if self.name == "Volume":
            return diag.min() <= 0
        if self.name == "Radius":
            return diag.min() <= self.p[0]
        return False

    def term(self, th, iso):
        import torch
        th = np.asarray(th, float)
        if self._outside(th, iso):
            return np.inf
        return float(self._value(torch.as_tensor(th, dtype=torch.float64), iso))

    def derivs(self, th, iso):
        import torch
        f = lambda t: self._value(t, iso)
        t = torch.as_tensor(np.asarray(th, float), dtype=torch.float64)
        g = torch.autograd.functional.jacobian(f, t)
        H = torch.autograd.functional.hessian(f, t)
        return g.numpy(), H.numpy()

    def total(self, I, x, r):
        iso = r.shape[1] == 1
        vals = [self.term(np.concatenate(([I[j]], x[j], r[j])), iso) for j in range(I.size)]
        return float(np.sum(vals)) if vals else 0.0


def _project_block(D, iso, t, th_old, d):
    """Projections of NewtonGaussian.linesearch_block (:159-187): block t of
    th = [I, x, r] moved by d, then I >= 0, x clamped to [-0.99, 0.99], r
    with the sign rule of GD_lib._step (the reference writes the 3D rule only)."""
    th = th_old.copy()
    o = (0, 1, 1 + D)[t]
    th[o:o + d.size] += d
    if t == 0:
        th[0] = max(0.0, th[0])
    elif t == 1:
        th[1:1 + D] = np.maximum(-0.99, np.minimum(0.99, th[1:1 + D]))
    else:
        _, _, th[1 + D:] = _project(D, iso, 0.0, np.zeros(D), th[1 + D:], np.zeros(th.size))
    return th


def newton_fit(I, x, r, geom, data, max_iter, tol=1e-4, min_iter=1, seed=1, reg=None, block=False,
               projector="point", n_sub=12, guess=None, trace=None):
    """NewtonGaussian.linesearch (block=False, :18-109, with GD_lib._step) or
    NewtonGaussian.linesearch_block (block=True, :112-237) with regulariser
    `reg` (a Reg, None = null), FP64.  Energies E = F + reg(all atoms),
    recomputed in full after every update as the reference does.

    Control flow (plain): h starts at 1 (:25); per update dx = solve(H + I/h,
    -g) (:68-74), the _step projections, then reject (E rises): if h > 2 eps
    the sweep is not converged, h /= 10; accept: if |dx| > tol the sweep is
    not converged, h *= 2; then h is clamped to [eps, 1/eps], eps = 1e-4
    (:81-92).  An inner loop of max_iter[1] updates per atom ends early when
    its last update left BREAK[1] set (:95-96); the fit ends after an outer
    iteration that left BREAK[0] set once its 0-based index exceeds min_iter
    (:97-98).  Block variant: blocks I, x, r in turn with h[t][j], x1.3 on
    accept (:205-217); the I block is skipped (no energy entry) unless
    ddf[0,0] > 1e-8 after damping, x and r when their solve fails (:165-187);
    the convergence test is |old - T| > tol |old| over the block (:214),
    where T is the unprojected I and x and the projected r (:170-199).

    The block variant restates the intent: its blocks are sliced with the
    spatial dimension (the reference slices with the string 'Ixr', :135,163,172,
    so that its x and r solves raise inside the try and are skipped).

    guess (callable (res, I1, x1, r1) -> (I1, x1, r1), res = data - R(atoms[:n])):
    greedy insertion (:38-50): for n = 1..N atom n-1 is set by guess, E's
    last entry is replaced by the energy of atoms[:n], and Newton sweeps run
    over atoms[:n]; with three entries in max_iter a last pass of
    max_iter[2] sweeps over all atoms follows.  Step sizes and the
    permutation stream run on across the stages.

    Returns (I, x, r, E, F): E after each update, F the fidelity of each
    update's proposal (also when rejected, as the reference records it).
    projector: 'point' (GaussTomo) or 'psf' (SingleParticleGaussTomo; the
    reference's Newton only calls Radon(.) and Radon.L2_derivs) or 'ft'
    (GaussFTVolume; geom is then the sample tuple (k0, k1, w0, w1)), or a
    (radproj, derivs) pair of callables.  trace (a list): one dict per
    update appended (atom j, block t, step size after the update, accept
    flag, the two BREAK flags after it), and {"j", "t", "skip": True} for a
    skipped block."""
    radproj, derivs = _projector(projector, n_sub)
    D = _atom_dim(projector, geom)
    I, x = np.array(I, float).reshape(-1), np.array(x, float).reshape(-1, D)
    r = np.array(r, float).reshape(I.size, -1)
    iso = r.shape[1] == 1
    n_all = I.size
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    max_iter = list(max_iter) + [1] * (2 - len(max_iter))
    eps = 1e-4
    nblk = 3 if block else 1
    h = [[1.0] * n_all for _ in range(nblk)]
    data = np.asarray(data)
    data = data.astype(complex) if np.iscomplexobj(data) else data.astype(float)
    rs = np.random.RandomState(seed)
    th_of = lambda j: np.concatenate(([I[j]], x[j], r[j]))

    def put(j, th):
        I[j], x[j], r[j] = th[0], th[1:1 + D], th[1 + D:]

    def total_reg(n):
        return reg.total(I[:n], x[:n], r[:n]) if reg is not None else 0.0

    E, F = [], []
    if guess is None:
        n = n_all
        R = radproj(I, x, r, geom)
        F.append(l2_squared_loss(R, data))
        E.append(F[0] + total_reg(n))
    else:
        n = 1
    while n <= n_all + 1:
        if n == n_all + 1:
            if guess is None or len(max_iter) == 2:
                break
            n = n_all
            max_iter = [max_iter[2], max_iter[1]]
        elif guess is not None:
            res = data - radproj(I[:n], x[:n], r[:n], geom)
            Ig, xg, rg = guess(res, I[n - 1:n].copy(), x[n - 1:n].copy(), r[n - 1:n].copy())
            I[n - 1], x[n - 1], r[n - 1] = np.reshape(Ig, -1)[0], np.reshape(xg, -1), np.reshape(rg, -1)
            R = radproj(I[:n], x[:n], r[:n], geom)
            Fv = l2_squared_loss(R, data)
            Ev = Fv + total_reg(n)
            if E:
                E[-1], F[-1] = Ev, Fv
            else:
                E, F = [Ev], [Fv]
        for it in range(max_iter[0]):
            brk = [True, True]
            for j in rs.permutation(np.arange(n)):
                for t in range(nblk):
                    for _ in range(max_iter[1]):
                        brk[1] = True
                        Rj = radproj(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom)
                        _, df, ddf = derivs(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom, data + Rj - R, "L2", 2)
                        df, ddf = df[0].copy(), ddf[0].copy()
                        old = th_of(j)
                        if reg is not None:
                            gr, Hr = reg.derivs(old, iso)
                            df += gr
                            ddf += Hr
                        ddf[np.diag_indices_from(ddf)] += 1.0 / h[t][j]
                        if not block:
                            dx = np.linalg.solve(ddf, -df)
                            I2, x2, r2 = _project(D, iso, old[0], old[1:1 + D], old[1 + D:], dx)
                            new = np.concatenate(([I2], x2, r2))
                        else:
                            o, m = (0, 1) if t == 0 else ((1, D) if t == 1 else (1 + D, old.size - 1 - D))
                            if t == 0:
                                if not ddf[0, 0] > 1e-8:
                                    if trace is not None:
                                        trace.append({"j": int(j), "t": t, "skip": True})
                                    break
                                d = np.array([-df[0] / ddf[0, 0]])
                            else:
                                try:
                                    d = np.linalg.solve(ddf[o:o + m, o:o + m], -df[o:o + m])
                                except np.linalg.LinAlgError:
                                    if trace is not None:
                                        trace.append({"j": int(j), "t": t, "skip": True})
                                    break
                            new = _project_block(D, iso, t, old, d)
                            # T of :170-199: the I and x blocks unprojected, r projected
                            moved_to = old[o:o + m] + d if t < 2 else new[o:o + m]
                        put(j, new)
                        Rnew = R + radproj(I[j:j + 1], x[j:j + 1], r[j:j + 1], geom) - Rj
                        Fn = l2_squared_loss(Rnew, data)
                        En = Fn + total_reg(n)
                        F.append(Fn)
                        acc = not En > E[-1]
                        if not acc:
                            put(j, old)
                            E.append(E[-1])
                            if h[t][j] > 2 * eps:
                                brk = [False, False]
                            h[t][j] /= 10
                        else:
                            R = Rnew
                            E.append(En)
                            if block:
                                moved = np.linalg.norm(old[o:o + m] - moved_to) > tol * np.linalg.norm(old[o:o + m])
                            else:
                                moved = np.sqrt(np.sum(dx * dx)) > tol
                            if moved:
                                brk = [False, False]
                            h[t][j] *= 1.3 if block else 2
                        h[t][j] = min(max(h[t][j], eps), 1 / eps)
                        if trace is not None:
                            trace.append({"j": int(j), "t": t, "h": h[t][j], "accept": acc, "brk": tuple(brk)})
                        if brk[1]:
                            break
            if brk[0] and it > min_iter:
                break
        n += 1
    return I, x, r, np.array(E), np.array(F)


# ---------------------------------------------------------------------------
# Fourier-space derivatives on hyperplane samples, GaussDictCode/bin/FTGaussian3D.py:55-92, 199-652
# ---------------------------------------------------------------------------
_FT_ROW = (0, 1, 2, 1, 2, 2)   # slot -> row of L in rk = L k (FTGaussian3D.py:231-234)
_FT_COL = (0, 1, 2, 0, 1, 0)   # slot -> column


def ft_plane_points(k0, k1, w0, w1=None):
    """Sample points of FTGaussian3D's hyperplane samplers: (nA, m0, m1, 3) with
    k = k0[i0] w0[a] + k1[i1] w1[a], or the 2.5D form (k0 w0[a,0], k0 w0[a,1], k1)."""
    k0, k1, w0 = (np.asarray(v, float) for v in (k0, k1, w0))
    if w1 is None:
        K = np.stack(np.broadcast_arrays(k0[None, :, None] * w0[:, None, None, 0],
                                         k0[None, :, None] * w0[:, None, None, 1],
                                         np.broadcast_to(k1[None, None, :], (w0.shape[0], k0.size, k1.size))), -1)
    else:
        w1 = np.asarray(w1, float)
        K = k0[None, :, None, None] * w0[:, None, None, :] + k1[None, None, :, None] * w1[:, None, None, :]
    return K


def ft_derivs(I, x, r, k0, k1, w0, w1, C, iso=False):
    """Per atom (each against all of C, as FTGaussian3D.derivs does for atom 0):
    f = 1/2 sum |g - C|^2, Df = Re sum (g - C) conj(dg), DDf = Re sum dg conj(dg)^T
    + (g - C) conj(ddg) over the hyperplane samples, g = I exp(-|L k|^2/2 - i x.k)
    (L lower-triangular from the slots, rk = L k; iso: |k|^2 / r^2), cut where
    |L k|^2 > 60 (FTGaussian3D.py:236-242, 456-462).  dg and ddg restate
    __derivs_aniso_aux / __derivs_iso_aux (:199-305, :436-497) per sample in
    FP64.  Returns f (n,), Df (n, P), DDf (n, P, P), P = 10 (5 iso)."""
    K = ft_plane_points(k0, k1, w0, w1).reshape(-1, 3)
    C = np.asarray(C, complex).reshape(-1)
    I, x = np.asarray(I, float).reshape(-1), np.asarray(x, float).reshape(-1, 3)
    r = np.asarray(r, float).reshape(I.size, -1)
    P = 5 if iso else 10
    f, Df, DDf = np.zeros(I.size), np.zeros((I.size, P)), np.zeros((I.size, P, P))
    for j in range(I.size):
        phase = np.exp(-1j * (K @ x[j]))
        if iso:
            rr = r[j, 0]
            N = (K * K).sum(1) / rr ** 2
            inside = N <= 60
            G = np.exp(-0.5 * N) * phase * inside
            g = I[j] * G
            dg = np.zeros((K.shape[0], 5), complex)
            ddg = np.zeros((K.shape[0], 5, 5), complex)
            dg[:, 0] = G
            dg[:, 1:4] = -1j * K * g[:, None]
            dg[:, 4] = N * g / rr
            ddg[:, 0, 1:4] = -1j * K * G[:, None]
            ddg[:, 0, 4] = N * G / rr
            ddg[:, 1:4, 1:4] = -K[:, :, None] * K[:, None, :] * g[:, None, None]
            ddg[:, 1:4, 4] = -1j * K * dg[:, 4:5]
            ddg[:, 4, 4] = (N - 3) * N * g / rr ** 2
        else:
            L = np.zeros((3, 3))
            for s in range(6):
                L[_FT_ROW[s], _FT_COL[s]] = r[j, s]
            rk = K @ L.T
            N = (rk * rk).sum(1)
            inside = N <= 60
            G = np.exp(-0.5 * N) * phase * inside
            g = I[j] * G
            dg = np.zeros((K.shape[0], 10), complex)
            ddg = np.zeros((K.shape[0], 10, 10), complex)
            dg[:, 0] = G
            dg[:, 1:4] = -1j * K * g[:, None]
            for s in range(6):
                a, b = _FT_ROW[s], _FT_COL[s]
                dg[:, 4 + s] = -rk[:, a] * K[:, b] * g
                ddg[:, 0, 4 + s] = -rk[:, a] * K[:, b] * G
                ddg[:, 1:4, 4 + s] = -1j * K * dg[:, 4 + s:5 + s]
                for t in range(6):
                    c, d = _FT_ROW[t], _FT_COL[t]
                    ddg[:, 4 + s, 4 + t] = g * (rk[:, a] * K[:, b] * rk[:, c] * K[:, d] - (a == c) * K[:, b] * K[:, d])
            ddg[:, 0, 1:4] = -1j * K * G[:, None]
            ddg[:, 1:4, 1:4] = -K[:, :, None] * K[:, None, :] * g[:, None, None]
        iu = np.triu_indices(P, 1)
        for i, l in zip(*iu):
            ddg[:, l, i] = ddg[:, i, l]
        res = g - C
        f[j] = 0.5 * np.sum(np.abs(res) ** 2)
        Df[j] = np.real(res @ np.conj(dg))
        DDf[j] = np.real(np.einsum("pi,pl->il", dg, np.conj(dg)) + np.einsum("p,pil->il", res, np.conj(ddg)))
    return f, Df, DDf


# ---------------------------------------------------------------------------
# Point-spread projector, GaussDictCode/bin/point_spread_functions.py:168-1350
# ---------------------------------------------------------------------------
def _cell_nodes(p, n_sub):
    """Gauss-Legendre nodes n_sub per cell of a uniform coordinate vector (cell
    = the interval of width |p[1] - p[0]| around each coordinate) and the
    weights of one cell."""
    p = np.asarray(p, float)
    hs = abs(p[1] - p[0]) / 2
    xg, wg = np.polynomial.legendre.leggauss(n_sub)
    return (p[:, None] + hs * xg[None, :]).reshape(-1), hs * wg


def _fine_geometry(geom, n_sub):
    D, T, lenT, w0, w1, p0, p1 = _geom_args(geom)
    f0, c0 = _cell_nodes(p0, n_sub)
    det = [f0]
    cw = [c0]
    if D == 3:
        f1, c1 = _cell_nodes(p1, n_sub)
        det.append(f1)
        cw.append(c1)
    return Geometry(geom.orientations, geom.ortho, det), cw


def psf_radproj(I, x, r, geom, n_sub=24):
    """Pixel-integrated projection P[a][k] = int_pixel R_a (RadProj,
    point_spread_functions.py:371-385, as the pixel integral it approximates):
    the pinned point projector radproj on n_sub Gauss-Legendre nodes per pixel
    and axis, FP64."""
    fine, cw = _fine_geometry(geom, n_sub)
    Rf = radproj(I, x, r, fine)
    nA = Rf.shape[0]
    if len(cw) == 1:
        return Rf.reshape(nA, -1, n_sub) @ cw[0]
    n0, n1 = Rf.shape[1] // n_sub, Rf.shape[2] // n_sub
    return np.einsum("aibjc,b,c->aij", Rf.reshape(nA, n0, n_sub, n1, n_sub), cw[0], cw[1])


def _psf_weights(C, cw, n_sub):
    """A pixel-valued array spread to the fine nodes with the node weights."""
    C = np.asarray(C, float)
    if len(cw) == 1:
        return (C[:, :, None] * cw[0][None, None, :]).reshape(C.shape[0], -1)
    W = C[:, :, None, :, None] * cw[0][None, None, :, None, None] * cw[1][None, None, None, None, :]
    return W.reshape(C.shape[0], C.shape[1] * n_sub, C.shape[2] * n_sub)


def psf_derivs(I, x, r, geom, C, mode="L2", order=2, n_sub=24, h=1e-6):
    """Per atom: mode 'linear' f = sum_k P_k C_k (derivs_RadProj,
    point_spread_functions.py:1314-1350), mode 'L2' f = 1/2 sum (P_k - C_k)^2
    (L2derivs_RadProj, :1077-1115), with gradient and Hessian.  Linear
    functionals of P are linear functionals of the line integral on the fine
    nodes (the pinned derivs, linear mode); the Gauss-Newton term of the L2
    Hessian takes the pixel gradients dP_k from central differences of
    psf_radproj (h relative)."""
    fine, cw = _fine_geometry(geom, n_sub)
    I, x = np.asarray(I, float).reshape(-1), np.asarray(x, float).reshape(I.size if np.ndim(I) else 1, -1)
    D = x.shape[1]
    r = np.asarray(r, float).reshape(I.size, -1)
    C = np.asarray(C, float)
    if mode == "linear":
        return derivs(I, x, r, fine, _psf_weights(C, cw, n_sub), "linear", order)
    P = 1 + D + r.shape[1]
    f = np.zeros(I.size)
    Df = np.zeros((I.size, P))
    DDf = np.zeros((I.size, P, P))
    for j in range(I.size):
        sl = slice(j, j + 1)
        Pj = psf_radproj(I[sl], x[sl], r[sl], geom, n_sub)
        res = Pj - C
        f[j] = 0.5 * np.sum(res * res)
        _, g, H = derivs(I[sl], x[sl], r[sl], fine, _psf_weights(res, cw, n_sub), "linear", order)
        Df[j] = g[0]
        if order > 1:
            th = np.concatenate(([I[j]], x[j], r[j]))
            J = []
            for i in range(P):
                e = np.zeros(P)
                e[i] = h * max(1.0, abs(th[i]))
                tp, tm = th + e, th - e
                Pp = psf_radproj(tp[:1], tp[1:1 + D][None], tp[1 + D:][None], geom, n_sub)
                Pm = psf_radproj(tm[:1], tm[1:1 + D][None], tm[1 + D:][None], geom, n_sub)
                J.append(((Pp - Pm) / (2 * e[i])).reshape(-1))
            J = np.array(J)
            DDf[j] = H[0] + J @ J.T
    return f, Df, DDf


def psf_volproj(I, x, r, coords, n_sub=12):
    """Voxel-integrated render u[v] = int_voxel f (VolProj,
    point_spread_functions.py:168-197): the pinned volproj on n_sub
    Gauss-Legendre nodes per voxel and axis, FP64."""
    nodes, cw = zip(*[_cell_nodes(c, n_sub) for c in coords])
    uf = volproj(I, x, r, list(nodes))
    shp = []
    for c in coords:
        shp += [len(c), n_sub]
    uf = uf.reshape(shp)
    if len(coords) == 2:
        return np.einsum("iajb,a,b->ij", uf, cw[0], cw[1])
    return np.einsum("iajbkc,a,b,c->ijk", uf, cw[0], cw[1], cw[2])


# ---------------------------------------------------------------------------
# Point-spread projector, the reference's own sampled rule
# (GaussDictCode/bin/point_spread_functions.py:25-88, 201-291, 389-527,
# 924-1075), FP64.  Values and derivatives come from one torch function of
# the atom's parameters so that the derivatives are those of the sampled sum
# itself (automatic differentiation with the samples held fixed).
# ---------------------------------------------------------------------------
def _psf_U_torch(rs, D):
    """U from the reference's slot order (:201-216, 389-404): 2D
    [[r0, r2], [0, r1]], 3D [[r0, r3, r5], [0, r1, r4], [0, 0, r2]]."""
    import torch
    z = torch.zeros((), dtype=rs.dtype)
    if D == 2:
        return torch.stack([torch.stack([rs[0], rs[2]]), torch.stack([z, rs[1]])])
    return torch.stack([torch.stack([rs[0], rs[3], rs[5]]), torch.stack([z, rs[1], rs[4]]),
                        torch.stack([z, z, rs[2]])])


def _psf_frames(geom):
    """Per angle (T, W0, W1) as the reference's kernels see them (:530-566,
    1096-1097, 1255-1263): 2D two-vectors and no W1; 2.5D T and W0 with a zero
    third entry and W1 = e_z; 3D as given."""
    D, T, lenT, w0, w1, p0, p1 = _geom_args(geom)
    out = []
    for a in range(T.shape[0]):
        if D == 2:
            out.append((T[a], w0[a], None))
        elif lenT == 2:
            out.append((np.array([T[a, 0], T[a, 1], 0.0]), np.array([w0[a, 0], w0[a, 1], 0.0]),
                        np.array([0.0, 0.0, 1.0])))
        else:
            out.append((T[a], w0[a], w1[a]))
    return out


def _psf_rule_axis(c, R, s):
    """Sample offsets from the cell centre and the axis weight of the
    reference's rule along one axis (:234-253, 440-458, 495-526): c = the
    atom's offset <x - y, W> from the cell centre, limits lo = max(c - R, -s)
    and hi = min(c + R, s), samples lo + k (hi - lo) / 10 for k = 1..11 (the
    first one step inside lo, the last one step beyond hi), each of weight
    (hi - lo) / 11 (the mean of the samples times hi - lo)."""
    lo = np.maximum(c - R, -s)
    hi = np.minimum(c + R, s)
    k = np.arange(1, 12, dtype=np.float64)
    return lo[..., None] + (0.1 * (hi - lo))[..., None] * k, (hi - lo) / 11.0


def _psf_atom_consts(tv, D):
    """R = 1 / min diag U (:25-30) and the atom's total, I (pi/30) / det U in
    3D and I sqrt(pi/30) / det U in 2D (:60-65), from the parameter values."""
    d = np.array([tv[1 + D + i] for i in range(D)])
    c = np.pi / 30 if D == 3 else np.sqrt(np.pi / 30)
    return 1.0 / d.min(), tv[0] * c / np.prod(d)


def _psf_sampled_atom(th, geom):
    """Sampled pixel values P[a, k0(, k1)] of one atom th = [I, x, U slots]
    (torch FP64 vector), the reference's __rf2 / __rf3 (:413-527) per pixel:
    n = distance of the projected centre from the pixel centre (__projnorm,
    :77-88); n + R < s gives the atom's total, n - R > s gives 0, otherwise
    the line integral I exp(-30 m) / |U T| (__Rf, :389-410, no cut-off) at
    the 11 (2D) or 11 x 11 samples of _psf_rule_axis, with s half the first
    detector axis' spacing for both axes.  Cases, samples and weights come
    from th's values and are held fixed, so derivatives are those of the
    integrand at the samples (__drf2 / __drf3, :924-1075; total pixels have
    none)."""
    import torch
    D, T, lenT, w0, w1, p0, p1 = _geom_args(geom)
    tv = th.detach().numpy()
    x = th[1:1 + D]
    U = _psf_U_torch(th[1 + D:], D)
    xv = tv[1:1 + D]
    R, total = _psf_atom_consts(tv, D)
    s = abs(p0[1] - p0[0]) / 2
    frames = _psf_frames(geom)
    nA = len(frames)
    npix = p0.size * (p1.size if D == 3 else 1)
    rows = []
    for a in range(nA):
        Ta, W0, W1 = frames[a]
        if D == 2:
            Y = p0[:, None] * W0[None, :]
        else:
            Y = (p0[:, None, None] * W0 + p1[None, :, None] * W1).reshape(-1, 3)
        X = xv[None, :] - Y
        t2 = (X * X).sum(1) - (X @ Ta) ** 2
        n = np.sqrt(np.where(t2 > 0, t2, 0.0))
        tot = n + R < s
        zero = ~tot & (n - R > s)
        samp = np.nonzero(~(tot | zero))[0]
        P = torch.zeros(npix, dtype=torch.float64)
        P = P.index_put((torch.from_numpy(np.nonzero(tot)[0]),), torch.full((int(tot.sum()),), total,
                                                                           dtype=torch.float64))
        if samp.size:
            Xs = X[samp]
            t0, wt = _psf_rule_axis(Xs @ W0, R, s)
            if D == 2:
                pts = Y[samp][:, None, :] + t0[:, :, None] * W0
            else:
                t1, wt1 = _psf_rule_axis(Xs @ W1, R, s)
                pts = (Y[samp][:, None, None, :] + t0[:, :, None, None] * W0 + t1[:, None, :, None] * W1)
                wt = wt * wt1
            Yt = torch.from_numpy(pts) - x
            rY = Yt @ U.T
            rT = U @ torch.from_numpy(Ta)
            inv = 1.0 / torch.sqrt((rT * rT).sum())
            rTh = rT * inv
            MY = rY - (rY @ rTh)[..., None] * rTh
            m = (MY * MY).sum(-1)
            v = th[0] * torch.exp(-30.0 * m) * inv
            vals = v.reshape(samp.size, -1).sum(1) * torch.from_numpy(wt)
            P = P.index_put((torch.from_numpy(samp),), vals)
        rows.append(P)
    shape = [nA, p0.size] + ([p1.size] if D == 3 else [])
    return torch.stack(rows).reshape(shape)


def _psf_theta(I, x, r, D):
    I, x = _c64(I).reshape(-1), _c64(x).reshape(-1, D)
    r = np.asarray(r, dtype=np.float64).reshape(I.size, -1)
    iso = r.shape[1] == 1
    return np.concatenate([I[:, None], x, _c64(to_aniso(r, D))], 1), iso


def psf_radproj_sampled(I, x, r, geom):
    """RadProj (point_spread_functions.py:371-385) with the reference's own
    sampled pixel rule (_psf_sampled_atom), summed over atoms, FP64."""
    import torch
    th, _ = _psf_theta(I, x, r, geom.dim)
    out = np.zeros(geom.shape)
    with torch.no_grad():
        for j in range(th.shape[0]):
            out += _psf_sampled_atom(torch.from_numpy(th[j]), geom).numpy()
    return out


def psf_derivs_sampled(I, x, r, geom, C, mode="L2", order=2):
    """Per atom (f, Df, DDf) of f = 1/2 sum (P - C)^2 (mode 'L2',
    L2derivs_RadProj :1077-1115) or f = sum P C ('linear', derivs_RadProj
    :1314-1350) with P the sampled rule: FP64 automatic differentiation of
    _psf_sampled_atom (samples held fixed).  Isotropic atoms are
    differentiated as U = rho Id and collapsed by the chain rule (the
    reference's own collapse, :1101-1113, drops terms)."""
    import torch
    D = geom.dim
    th, iso = _psf_theta(I, x, r, D)
    Ct = torch.from_numpy(_c64(C).reshape(geom.shape))
    N, P = th.shape
    f = np.zeros(N)
    Df = np.zeros((N, P))
    DDf = np.zeros((N, P, P))

    def fun(t):
        Pj = _psf_sampled_atom(t, geom)
        if mode == "L2":
            d = Pj - Ct
            return 0.5 * (d * d).sum()
        return (Pj * Ct).sum()

    for j in range(N):
        t = torch.from_numpy(th[j].copy()).requires_grad_(True)
        val = fun(t)
        f[j] = val.item()
        if not val.requires_grad:
            continue          # no sampled pixel: total and zero pixels have no derivative
        (g,) = torch.autograd.grad(val, t)
        Df[j] = g.numpy()
        if order > 1:
            DDf[j] = torch.autograd.functional.hessian(fun, torch.from_numpy(th[j].copy())).numpy()
    if iso:
        E = iso_collapse(D)
        Df = Df @ E
        DDf = np.einsum("ai,nab,bj->nij", E, DDf, E)
    return f, Df, DDf


def psf_volproj_sampled(I, x, r, coords):
    """VolProj (point_spread_functions.py:168-197) with the reference's own
    sampled voxel rule (__if2 / __if3, :222-291), FP64: per (atom, voxel) n =
    |x - y|; n + R < s gives the atom's total, n - R > s gives 0, otherwise
    I sqrt(30/pi) exp(-30 |U (y' - x)|^2) (__f, :201-219) at the 11^D samples
    y' of _psf_rule_axis per axis, with s half the second axis' spacing for
    every axis (:302, 324, 347, 362)."""
    D = len(coords)
    y = [_c64(c).reshape(-1) for c in coords]
    th, _ = _psf_theta(I, x, r, D)
    s = abs(y[1][1] - y[1][0]) / 2
    grids = np.meshgrid(*y, indexing="ij")
    Yv = np.stack([g.reshape(-1) for g in grids], 1)
    out = np.zeros(Yv.shape[0])
    for j in range(th.shape[0]):
        tv = th[j]
        xv = tv[1:1 + D]
        R, total = _psf_atom_consts(tv, D)
        Um = np.zeros((D, D))
        rs = tv[1 + D:]
        if D == 2:
            Um[0, 0], Um[1, 1], Um[0, 1] = rs[0], rs[1], rs[2]
        else:
            Um[0, 0], Um[1, 1], Um[2, 2], Um[0, 1], Um[1, 2], Um[0, 2] = rs
        X = xv[None, :] - Yv
        n = np.sqrt((X * X).sum(1))
        tot = n + R < s
        zero = ~tot & (n - R > s)
        samp = np.nonzero(~(tot | zero))[0]
        out[tot] += total
        if samp.size:
            offs, wts = zip(*[_psf_rule_axis(X[samp, i], R, s) for i in range(D)])
            if D == 2:
                pts = np.stack(np.broadcast_arrays(Yv[samp, 0, None, None] + offs[0][:, :, None],
                                                   Yv[samp, 1, None, None] + offs[1][:, None, :]), -1)
            else:
                pts = np.stack(np.broadcast_arrays(Yv[samp, 0, None, None, None] + offs[0][:, :, None, None],
                                                   Yv[samp, 1, None, None, None] + offs[1][:, None, :, None],
                                                   Yv[samp, 2, None, None, None] + offs[2][:, None, None, :]), -1)
            d = (pts - xv) @ Um.T
            v = tv[0] * np.sqrt(30 / np.pi) * np.exp(-30 * (d * d).sum(-1))
            out[samp] += v.reshape(samp.size, -1).sum(1) * np.prod(np.stack(wts, 1), 1)
    return out.reshape(tuple(len(c) for c in y))


# ---------------------------------------------------------------------------
# KL moment initialiser, KL_GaussRadon.py:57-394
# ---------------------------------------------------------------------------
def kl_sino_moments(f, geom):
    """Per angle: sum f, sum f p0, sum f p1, sum f p0^2, sum f p1^2 over the
    detector (the sums MOM0 / MOM1 / MOM2 of KL_GaussRadon.py:221-392 are
    made of; the reference's __MOM1_2D runs its inner loop over shape[1]
    instead of shape[2], :273 — restated over the whole detector)."""
    f = np.asarray(f, float)
    p0 = np.asarray(geom.detector[0], float)
    nA = f.shape[0]
    f3 = f.reshape(nA, p0.size, -1)
    p1 = np.asarray(geom.detector[1], float) if geom.dim == 3 else np.zeros(1)
    rows = f3.sum(2)   # (nA, n0)
    cols = f3.sum(1)   # (nA, n1)
    return np.stack([rows.sum(1), rows @ p0, cols @ p1, rows @ (p0 * p0), cols @ (p1 * p1)], 1)


def kl_newton_sino(a, b, t):
    """f with f + t log f = a + t log b per pixel (Newton_sino_2D,
    KL_GaussRadon.py:193-229): f = a when t == 0, f = 0 where b <= 1e-15;
    else the root, by bisection on u = log f of the increasing
    h(u) = e^u + t u over a bracket that contains it (an independent method:
    the reference iterates Newton ten times).  Roots below the FP64 range
    come out as 0, as the reference's underflow gives."""
    a, b = np.asarray(a, float), np.asarray(b, float)
    if t == 0:
        return a.copy()
    c0 = a + t * np.log(np.where(b > 1e-15, b, 1.0))
    # h(lo) <= c0: lo <= 0 gives e^lo <= 1, and t lo <= c0 - 1 - t
    lo = np.minimum((c0 - 1.0) / t - 1.0, 0.0)
    # h(hi) >= c0: e^hi >= |c0| + 1 and t hi >= 0
    hi = np.log(np.abs(c0) + 1.0) + 1.0
    for _ in range(300):
        mid = 0.5 * (lo + hi)
        below = np.exp(mid) + t * mid < c0
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return np.where(b > 1e-15, np.exp(0.5 * (lo + hi)), 0.0)


def kl_orthog_op(If, Ify, T):
    """op = sum_a If[a] (Id - T_a T_a^T), vec = sum_a (Id - T_a T_a^T) Ify[a]
    (orthog_op, KL_GaussRadon.py:256-331).  2.5D (2-vectors T, 3-vectors
    Ify): the third axis is the rotation axis, untouched by the projector."""
    If, Ify, T = np.asarray(If, float), np.asarray(Ify, float), np.asarray(T, float)
    d = Ify.shape[1]
    op, vec = np.zeros((d, d)), np.zeros(d)
    k = T.shape[1]
    for a in range(T.shape[0]):
        P = np.eye(d)
        P[:k, :k] -= np.outer(T[a], T[a])
        op += If[a] * P
        vec += P @ Ify[a]
    return op, vec


def kl_step(S0, I, x, r, geom, t, lagrange=False, projector="point", n_sub=12, max_iter=100):
    """One atom from a residual by moment matching (doKL_ProjGDStep_iso,
    KL_GaussRadon.py:57-134; lagrange=True: doKL_LagrangeStep_iso, :137-190),
    FP64.  S0 is the residual with the atom's own projection added back
    (res + R(atom), :74).  The atom is reset (I = 1, x = 0, diagonal slots
    10, others 0), then per iteration: centre M from the first moments
    (orthog_op), spread sigma^2 = sum_a int f |P_a (y - M)|^2 / ((D - 1) int f),
    radius slots 1 / (sigma sqrt 60) so that the projection exp(-30 m) has
    that spread, intensity by the mass ratio int f / int R(atom); then f is
    updated from S0 and the new projection.  Stops when sigma, M and I change
    by less than 1e-4 (relative).

    Restated intent where the reference's code differs: the reference sets
    iso r = sigma and aniso r = 1 / sigma (:111-114) — neither is the radius
    of an exp(-30 m) atom of spread sigma; its r[0, :3] = 10 writes the 2D
    off-diagonal slot (:83); with no positive mass it narrows the atom by
    halving its width until below 1e-4 (the iso branch of :92-96; the aniso
    branch widens it)."""
    radproj_fn, _ = _projector(projector, n_sub)
    D = geom.dim
    T = np.asarray(geom.orientations, float)
    lenT = T.shape[1]
    w0 = np.asarray(geom.ortho[0], float)
    w1 = np.asarray(geom.ortho[1], float) if (D == 3 and lenT == 3) else None
    S0 = np.asarray(S0, float)
    iso = np.asarray(r).reshape(1, -1).shape[1] == 1
    I = np.array([1.0])
    x = np.zeros((1, D))
    r = np.zeros((1, 1 if iso else D * (D + 1) // 2))
    r[0, :1 if iso else D] = 10.0
    sig = 1.0 / (10.0 * np.sqrt(60.0))
    f = np.maximum(0.0, S0)
    for _ in range(max_iter):
        m = kl_sino_moments(f, geom)
        If = m[:, 0]
        if If.sum() < 1e-8:
            if sig < 1e-4:
                break
            sig *= 0.5
            r *= 2.0
            continue
        if D == 2:
            Ify = m[:, 1:2] * w0
        elif lenT == 2:
            Ify = np.concatenate((m[:, 1:2] * w0, m[:, 2:3]), 1)
        else:
            Ify = m[:, 1:2] * w0 + m[:, 2:3] * w1
        op, vec = kl_orthog_op(If, Ify, T)
        try:
            M = np.linalg.solve(op, vec)
        except np.linalg.LinAlgError:
            M = np.zeros_like(vec)
        c0 = w0 @ M[:w0.shape[1]]
        Ifyy = np.sum(m[:, 3] - 2 * c0 * m[:, 1] + c0 * c0 * m[:, 0])
        if D == 3:
            c1 = w1 @ M if w1 is not None else np.full(T.shape[0], M[2])
            Ifyy += np.sum(m[:, 4] - 2 * c1 * m[:, 2] + c1 * c1 * m[:, 0])
        old = (sig, x[0].copy(), I[0])
        x[0] = M
        sig = np.sqrt(max(Ifyy, 0.0) / ((D - 1) * If.sum()))
        if not sig > 0:
            break
        r[0, :1 if iso else D] = 1.0 / (sig * np.sqrt(60.0))
        RR = radproj_fn(I, x, r, geom)
        tot = RR.sum()
        if not tot > 0:
            break
        A = If.sum() * I[0] / tot
        RR *= A / I[0]
        I[0] = A
        moved = (abs(old[0] - sig) > 1e-4 * abs(old[0]) or np.linalg.norm(old[1] - M) > 1e-4 * np.linalg.norm(old[1])
                 or abs(old[2] - A) > 1e-4 * abs(old[2]))
        if not moved:
            break
        f = kl_newton_sino(S0, RR, t) if lagrange else np.maximum(0.0, t * S0 + (1 - t) * RR)
    return I, x, r
