"""KL moment initialiser over the HIP library (KL_GaussRadon.py:57-394).

    sino_moments(S, space, Q=None, mode, a, b)   per-angle FP64 moments of f
                                                 (gt_sino_moments)
    doKL_ProjGDStep_iso(res, atom, t, R)         KL_GaussRadon.py:57-134
    doKL_LagrangeStep_iso(res, atom, t, R)       KL_GaussRadon.py:137-190
    orthog_op(If, Ify, T)                        KL_GaussRadon.py:256-331

The initialisers place one atom on a residual sinogram by matching moments:
with f >= 0 the part of the residual to explain, the atom's centre M solves
sum_a If_a (Id - T_a T_a^T) M = sum_a (Id - T_a T_a^T) Ify_a (first moments),
its spread is sigma^2 = sum_a int f |P_a(y - M)|^2 / ((D - 1) sum_a If_a), its
radius slots 1 / (sigma sqrt 60) (the projection of exp(-30 |U(y - x)|^2)
has spread 1 / (60 rho^2) per detector axis) and its intensity the mass
ratio int f / int R(atom).  f is then re-formed from the residual and the
new projection, and the loop runs until sigma, M and I settle (1e-4
relative, at most 100 rounds).  All sinogram passes are one kernel each; the
3x3 solve is on the host.

Where the reference's code and its stated mathematics differ, the
mathematics is followed: its iso radius is set to sigma and the aniso
diagonal to 1 / sigma (KL_GaussRadon.py:111-114), neither the radius of an
exp(-30 m) atom of spread sigma; its r[0, :3] = 10 also writes the 2D
off-diagonal slot (:83); its MOM1 sums only the first n1 = n0 columns
(:273); with no positive mass left the atom is narrowed (width halved) until
its width is below 1e-4 and returned, for both layouts (the reference
narrows iso atoms and widens aniso ones in the ProjGD step, :92-96).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import ops
from ._lib import GT_KL_CLAMP, GT_KL_LAGRANGE, check, load, stream_handle

_SQRT60 = float(np.sqrt(60.0))


def sino_moments(S, space, Q=None, mode=GT_KL_CLAMP, a=1.0, b=0.0):
    """(nA, 5) FP64 host array: per angle sum f, sum f p0, sum f p1,
    sum f p0^2, sum f p1^2 with f = max(0, a S + b Q) (GT_KL_CLAMP) or the
    root of f + a log f = S + a log(b Q) (GT_KL_LAGRANGE)."""
    return sino_moments_device(S, space, Q, mode, a, b).cpu().numpy()


def sino_moments_device(S, space, Q=None, mode=GT_KL_CLAMP, a=1.0, b=0.0, out=None):
    """sino_moments into a device (nA, 5) FP64 tensor, enqueued on the
    current stream without waiting."""
    lib = load()
    dev = S.device
    geom = ops.device_geometry(space, dev)
    S = S.to(torch.float32).contiguous()
    if tuple(S.shape) != tuple(geom.shape):
        raise ValueError(f"S must have shape {geom.shape}, got {tuple(S.shape)}")
    if Q is not None:
        Q = Q.to(torch.float32).contiguous()
        if tuple(Q.shape) != tuple(geom.shape):
            raise ValueError(f"Q must have shape {geom.shape}, got {tuple(Q.shape)}")
    if out is None:
        out = torch.empty((geom.nA, 5), dtype=torch.float64, device=dev)
    if tuple(out.shape) != (geom.nA, 5) or out.dtype != torch.float64 or not out.is_contiguous():
        raise ValueError(f"out must be a contiguous float64 tensor of shape ({geom.nA}, 5)")
    need = lib.gt_sino_moments_workspace(ctypes.byref(geom.struct))
    ws = ops._ws.get(need, dev)
    check(lib.gt_sino_moments(ctypes.byref(geom.struct), ops._ptr(S), ops._ptr(Q), int(mode), float(a), float(b),
                              ops._ptr(out), ops._ptr(ws), ws.numel(), stream_handle(dev)), "gt_sino_moments")
    return out


def orthog_op(If, Ify, T):
    """op = sum_a If[a] (Id - T_a T_a^T), vec = sum_a (Id - T_a T_a^T) Ify[a];
    with 2-vectors T and 3-vectors Ify (2.5D) the third axis is the rotation
    axis, which the projection leaves alone."""
    If, Ify, T = np.asarray(If, float), np.asarray(Ify, float), np.asarray(T, float)
    d, k = Ify.shape[1], T.shape[1]
    P = np.broadcast_to(np.eye(d), (T.shape[0], d, d)).copy()
    P[:, :k, :k] -= T[:, :, None] * T[:, None, :]
    return np.einsum("a,aij->ij", If, P), np.einsum("aij,aj->i", P, Ify)


def _project(R, atom):
    out = R(atom)
    return out.data if hasattr(out, "data") and not isinstance(out, torch.Tensor) else out


def _kl_step(res, atom, t, R, lagrange):
    space = R.range
    dev = atom.I.device
    D = atom.x.shape[1]
    iso = atom.r.shape[1] == 1
    res = res.data if hasattr(res, "data") and not isinstance(res, torch.Tensor) else res
    res = torch.as_tensor(res, device=dev).to(torch.float32)
    S0 = (res + _project(R, atom)).contiguous()
    T = np.asarray(space.orientations, float)
    w0 = np.asarray(space.ortho[0], float)
    w1 = np.asarray(space.ortho[1], float) if (D == 3 and T.shape[1] == 3) else None

    def setr(rho):
        r = torch.zeros_like(atom.r)
        r[0, :1 if iso else D] = float(rho)
        atom.r.copy_(r)

    atom.I.fill_(1.0)
    atom.x.zero_()
    rho = 10.0
    setr(rho)
    sig = 1.0 / (rho * _SQRT60)
    I = 1.0
    x = np.zeros(D)
    mode, a, b, Q = GT_KL_CLAMP, 1.0, 0.0, None   # f = max(0, S0)
    for _ in range(100):
        m = sino_moments(S0, space, Q, mode, a, b)
        If = m[:, 0]
        if If.sum() < 1e-8:
            if sig < 1e-4:
                break
            sig *= 0.5
            rho *= 2.0
            setr(rho)
            continue
        if D == 2:
            Ify = m[:, 1:2] * w0
        elif w1 is None:
            Ify = np.concatenate((m[:, 1:2] * w0, m[:, 2:3]), 1)
        else:
            Ify = m[:, 1:2] * w0 + m[:, 2:3] * w1
        op, vec = orthog_op(If, Ify, T)
        try:
            M = np.linalg.solve(op, vec)
        except np.linalg.LinAlgError:
            M = np.zeros_like(vec)
        c0 = w0 @ M[:w0.shape[1]]
        Ifyy = np.sum(m[:, 3] - 2 * c0 * m[:, 1] + c0 * c0 * m[:, 0])
        if D == 3:
            c1 = w1 @ M if w1 is not None else np.full(T.shape[0], M[2])
            Ifyy += np.sum(m[:, 4] - 2 * c1 * m[:, 2] + c1 * c1 * m[:, 0])
        old = (sig, x.copy(), I)
        x = M
        atom.x.copy_(torch.as_tensor(M, dtype=torch.float32, device=dev).reshape(1, D))
        sig = float(np.sqrt(max(Ifyy, 0.0) / ((D - 1) * If.sum())))
        if not sig > 0:  # all mass on the centre's pixels: no spread to match
            break
        rho = 1.0 / (sig * _SQRT60)
        setr(rho)
        RR = _project(R, atom).contiguous()
        tot = float(sino_moments(RR, space)[:, 0].sum())
        if not tot > 0:
            break
        A = If.sum() * I / tot
        scale = A / I
        I = A
        atom.I.fill_(A)
        moved = (abs(old[0] - sig) > 1e-4 * abs(old[0]) or np.linalg.norm(old[1] - M) > 1e-4 * np.linalg.norm(old[1])
                 or abs(old[2] - A) > 1e-4 * abs(old[2]))
        if not moved:
            break
        # next f from S0 and the projection of the new atom (RR scaled to intensity A)
        if lagrange:
            mode, a, b, Q = GT_KL_LAGRANGE, float(t), scale, RR
        else:
            mode, a, b, Q = GT_KL_CLAMP, float(t), (1.0 - float(t)) * scale, RR
    return atom


def doKL_ProjGDStep_iso(res, atom, t, R):
    """KL_GaussRadon.py:57-134.  res = data - R(recon[:n]) with `atom` among
    recon[:n]; fits `atom` (one atom, modified in place and returned) to
    res + R(atom) by moment matching, with f = max(0, t S0 + (1 - t) R(atom))
    between rounds.  R: the projector operator (GaussTomo or
    SingleParticleGaussTomo)."""
    return _kl_step(res, atom, t, R, False)


def doKL_LagrangeStep_iso(res, atom, t, R):
    """KL_GaussRadon.py:137-190: as doKL_ProjGDStep_iso with f the
    per-pixel root of f + t log f = S0 + t log R(atom) between rounds."""
    return _kl_step(res, atom, t, R, True)


def KL_2D(f, g):
    """sum f (log(f / g) - 1) + g over the entries where it is defined
    (KL_GaussRadon.py:51-54, nansum)."""
    f = torch.as_tensor(f).double()
    g = torch.as_tensor(g, device=f.device).double()
    v = f * (torch.log(f / g) - 1) + g
    return float(torch.nansum(v))
