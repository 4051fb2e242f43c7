"""Fitters over the HIP library.

newton_linesearch        NewtonGaussian.linesearch (NewtonGaussian.py:18-109)
                         with GD_lib._step (GD_lib.py:492-536): device-resident,
                         a fit advanced by a group of workgroups that share its
                         angles (gt_fit_run_blocks), FP64 solves and FP64 energy
                         deltas.
newton_linesearch_block  NewtonGaussian.linesearch_block (NewtonGaussian.py:
                         112-237): the same kernel updating I, x and r of each
                         atom in turn.
gd_linesearch_block      GD_lib.linesearch_block (GD_lib.py:82-178): block
                         gradient descent; device-resident for the point
                         projector (the same kernel, block = 2), driven from
                         the host over the operator's kernels otherwise.
newton_host              the Newton fitters driven from the host over an
                         operator's forward and derivative kernels: what the
                         drop-ins run for SingleParticleGaussTomo, whose
                         pixel-integrated projector the device fitter does not
                         implement (NewtonGaussian.linesearch only calls
                         Radon(.) and Radon.L2_derivs, so it fits with any
                         operator).

The fidelity is l2 (transport_loss.py:567-600); the Newton fitters take any
regulariser of the regularisation module, evaluated inside the fit kernel.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import ops, psf
from ._lib import check, load, stream_handle
from .regularisation import as_regulariser


class gt_fit_slot(ctypes.Structure):
    _fields_ = [
        ("active", ctypes.c_int), ("iso", ctypes.c_int), ("N", ctypes.c_int), ("nA", ctypes.c_int),
        ("n0", ctypes.c_int), ("n1", ctypes.c_int), ("lenT", ctypes.c_int), ("sorted", ctypes.c_int),
        ("max_outer", ctypes.c_int), ("max_inner", ctypes.c_int), ("min_iter", ctypes.c_int), ("quota", ctypes.c_int),
        ("tol", ctypes.c_double),
        ("T", ctypes.c_void_p), ("w0", ctypes.c_void_p), ("w1", ctypes.c_void_p), ("p0", ctypes.c_void_p),
        ("p1", ctypes.c_void_p), ("perm", ctypes.c_void_p),
        ("I", ctypes.c_void_p), ("x", ctypes.c_void_p), ("r", ctypes.c_void_p), ("resid", ctypes.c_void_p),
        ("h", ctypes.c_void_p), ("Ehist", ctypes.c_void_p), ("scratch", ctypes.c_void_p), ("stage", ctypes.c_void_p),
        ("it", ctypes.c_int), ("pos", ctypes.c_int), ("inner", ctypes.c_int), ("jj", ctypes.c_int),
        ("brk0", ctypes.c_int), ("brk1", ctypes.c_int), ("done", ctypes.c_int), ("pad1", ctypes.c_int),
        ("E", ctypes.c_double), ("updates", ctypes.c_longlong), ("accepted", ctypes.c_longlong),
        ("phase_ticks", ctypes.c_void_p),
        ("block", ctypes.c_int), ("reg_kind", ctypes.c_int), ("reg_p", ctypes.c_double * 5),
        ("Fhist", ctypes.c_void_p),
        ("blk", ctypes.c_int), ("ninf", ctypes.c_int), ("F", ctypes.c_double), ("Rsum", ctypes.c_double),
        ("gd_nblk", ctypes.c_int), ("gd_blocks", ctypes.c_int * 3),
    ]


SLOT_BYTES = ctypes.sizeof(gt_fit_slot)
MAX_ANGLES = 256
MAX_DETECTOR = 1024


class gt_fit_block(ctypes.Structure):
    _fields_ = [("slot", ctypes.c_int), ("member", ctypes.c_int), ("group", ctypes.c_int), ("quota", ctypes.c_int)]


BLOCK_BYTES = ctypes.sizeof(gt_fit_block)
ANGLES_PER_MEMBER = 16  # one angle per wave of a 16-wave workgroup


def _declare():
    return load()


def default_group(nA):
    """Workgroups that advance one fit together: enough that each holds about
    one angle per wave, at most 8."""
    return max(1, min(8, -(-int(nA) // ANGLES_PER_MEMBER)))


def max_blocks():
    """Fit workgroups the device holds at once (grouped launches may not exceed it)."""
    n = _declare().gt_fit_max_blocks()
    if n <= 0:
        check(-1, "gt_fit_max_blocks")
    return n


def block_table(groups, quotas=None):
    """gt_fit_block entries for slots 0.. with the given group sizes (None or 0
    leaves a slot out): members of one slot are spread over consecutive blocks."""
    ent = []
    for s, g in enumerate(groups):
        if not g:
            continue
        q = 0 if quotas is None else int(quotas[s])
        ent += [(s, m, int(g), q) for m in range(int(g))]
    return np.array(ent, dtype=np.int32).reshape(-1, 4)


def permutations(n, max_outer, seed=1):
    """Atom order of every outer iteration: random.seed(1) then one
    random.permutation(arange(n)) per outer iteration (NewtonGaussian.py:37, 55)."""
    rs = np.random.RandomState(seed)
    return np.stack([rs.permutation(np.arange(n)) for _ in range(max_outer)]).astype(np.int32)


class FitRequest:
    """Device state of one Newton fit: atoms, residual, step sizes, history.

    reg: a regularisation.Regulariser (None: null); block: update I, x and r
    of each atom in turn (NewtonGaussian.linesearch_block); gd_dims (a
    string over "xrI" in the reference's dims order): block gradient descent
    over those blocks instead (GD_lib.linesearch_block); perm: the atom
    order of each outer iteration (default: permutations(N, max_outer,
    seed)); h: starting step sizes ([blocks][N]; default 1, 1e-2 for gradient
    descent); stage: give the fit a staging slab (the energy pass then
    commits at once and a rejected update is undone from the slab; without
    it an accepted update is committed by a second pass).  Same results
    either way; off by default: the second pass re-reads what the energy
    pass has just brought into cache, while the slab doubles the energy
    pass's stores and the undo is a second full copy (measured on the
    serving workload's shapes, 5-47% less time per update without it, one
    fit alone or 256 at once)."""

    def __init__(self, atoms, space, data, max_iter, tol=1e-4, min_iter=1, seed=1, record=True, device=None,
                 reg=None, block=False, perm=None, h=None, stage=False, gd_dims=None):
        dev = device or atoms.I.device
        self.reg = as_regulariser(reg)
        # variant: 0 Newton, 1 block Newton, 2 block gradient descent over gd_dims
        self.block = bool(block)
        self.gd_blocks = [] if gd_dims is None else [{"I": 0, "x": 1, "r": 2}[c] for c in gd_dims]
        if gd_dims is not None and not 1 <= len(self.gd_blocks) <= 3:
            raise ValueError("gd_dims must name one to three of the blocks x, r, I")
        self.mode = 2 if gd_dims is not None else int(self.block)
        nblk = len(self.gd_blocks) if self.mode == 2 else (3 if self.block else 1)
        self.nblk = nblk
        if not hasattr(max_iter, "__iter__"):
            max_iter = [max_iter, 1]
        max_iter = list(max_iter) + [1] * (2 - len(max_iter))
        self.max_outer, self.max_inner = int(max_iter[0]), int(max_iter[1])
        self.geom = ops.device_geometry(space, dev)
        if self.geom.nA > MAX_ANGLES:
            raise ValueError(f"the fitter supports at most {MAX_ANGLES} angles, got {self.geom.nA}")
        if max(self.geom.n0, self.geom.n1) > MAX_DETECTOR:
            raise ValueError(f"the fitter supports detectors up to {MAX_DETECTOR} pixels per axis")
        self.dim = self.geom.dim
        self.I = atoms.I.detach().to(dev, torch.float32).clone().contiguous()
        self.x = atoms.x.detach().to(dev, torch.float32).clone().contiguous()
        self.r = atoms.r.detach().to(dev, torch.float32).clone().contiguous()
        self.N = int(self.I.numel())
        self.iso = int(self.r.shape[1] == 1)
        data = data.data if hasattr(data, "data") and not isinstance(data, torch.Tensor) else data
        self.data = torch.as_tensor(data, device=dev).to(torch.float32).reshape(self.geom.shape).contiguous()
        self.resid = ops.forward(self, space).sub_(self.data)
        r64 = self.resid.double()
        self.F0 = float(0.5 * (r64 * r64).sum())
        # E = F + reg(recon) (NewtonGaussian.py:28-30): the sum of the finite
        # per-atom terms and the count of the others
        self.ninf0, self.Rsum0 = 0, 0.0
        if self.reg.kind != 0 and self.N > 0:
            val = self.reg.terms(self, 0)[0]
            fin = torch.isfinite(val)
            self.ninf0 = int((~fin).sum())
            self.Rsum0 = float(val[fin].sum())
        self.E0 = self.F0 + (float("inf") if self.ninf0 else self.Rsum0)
        if perm is None:
            perm = permutations(self.N, self.max_outer, seed)
        perm = np.asarray(perm, dtype=np.int32).reshape(-1, self.N)
        if perm.shape[0] < self.max_outer or not all(np.array_equal(np.sort(p), np.arange(self.N)) for p in perm):
            raise ValueError("perm must hold max_outer permutations of range(N)")
        self.perm = torch.as_tensor(perm[: self.max_outer].copy(), device=dev).contiguous()
        if h is None:
            h0 = 1e-2 if self.mode == 2 else 1.0  # GD_lib.py:99, NewtonGaussian.py:25,119
            self.h = torch.full((nblk * self.N,), h0, dtype=torch.float64, device=dev)
        else:
            h = np.asarray(h, dtype=np.float64).reshape(-1)
            if h.size != nblk * self.N:
                raise ValueError(f"h must hold {nblk * self.N} step sizes")
            self.h = torch.as_tensor(h.copy(), device=dev)
        nh = nblk * self.max_outer * self.max_inner * self.N + 1
        self.Ehist = torch.zeros(nh, dtype=torch.float64, device=dev) if record else None
        self.Fhist = torch.zeros(nh, dtype=torch.float64, device=dev) if record else None
        if record:
            self.Ehist[0] = self.E0
            self.Fhist[0] = self.F0
        lib = _declare()
        self.scratch = torch.zeros(lib.gt_fit_scratch_bytes(self.geom.nA), dtype=torch.uint8, device=dev)
        # old residual of the update in flight: lets the energy pass commit at once
        self.stage = torch.empty(self.geom.shape, dtype=torch.float32, device=dev) if stage else None
        g = self.geom
        self.slot = gt_fit_slot(
            active=1 if self.N > 0 else 0, iso=self.iso, N=self.N, nA=g.nA, n0=g.n0, n1=g.n1, lenT=g.lenT,
            sorted=g.sorted, max_outer=self.max_outer, max_inner=self.max_inner, min_iter=int(min_iter), quota=0,
            tol=float(tol), T=g.T.data_ptr(), w0=g.w0.data_ptr(), w1=g.w1.data_ptr() if g.w1 is not None else None,
            p0=g.p0.data_ptr(), p1=g.p1.data_ptr() if g.p1 is not None else None, perm=self.perm.data_ptr(),
            I=self.I.data_ptr(), x=self.x.data_ptr(), r=self.r.data_ptr(), resid=self.resid.data_ptr(),
            h=self.h.data_ptr(), Ehist=self.Ehist.data_ptr() if record else None, scratch=self.scratch.data_ptr(),
            stage=self.stage.data_ptr() if stage else None,
            it=0, pos=0, inner=0, jj=0, brk0=1, brk1=1, done=0 if self.N > 0 else 1, pad1=0, E=self.E0,
            updates=0, accepted=0, block=self.mode, reg_kind=int(self.reg.kind),
            Fhist=self.Fhist.data_ptr() if record else None, blk=0, ninf=self.ninf0, F=self.F0, Rsum=self.Rsum0,
            gd_nblk=len(self.gd_blocks))
        self.slot.reg_p[:] = [float(v) for v in self.reg.params()]
        for i, b in enumerate(self.gd_blocks):
            self.slot.gd_blocks[i] = b

    def slot_bytes(self):
        return bytes(self.slot)


def read_slots(dev_table, n):
    host = dev_table[: n * SLOT_BYTES].cpu().numpy().tobytes()
    return [gt_fit_slot.from_buffer_copy(host, i * SLOT_BYTES) for i in range(n)]


def run_slots(dev_table, nslots, dim, budget):
    lib = _declare()
    check(lib.gt_fit_run(ctypes.c_void_p(dev_table.data_ptr()), nslots, dim, budget, stream_handle(dev_table.device)),
          "gt_fit_run")


def run_blocks(dev_table, nslots, dev_blocks, nblocks, dim, budget):
    """Grouped launch: dev_blocks is a device int32 tensor of gt_fit_block rows."""
    lib = _declare()
    check(lib.gt_fit_run_blocks(ctypes.c_void_p(dev_table.data_ptr()), nslots, ctypes.c_void_p(dev_blocks.data_ptr()),
                                nblocks, dim, budget, stream_handle(dev_table.device)), "gt_fit_run_blocks")


def newton_fit(atoms, space, data, max_iter, tol=1e-4, min_iter=1, seed=1, budget=None, record=True, group=None,
               reg=None, block=False, perm=None, h=None, gd_dims=None):
    """Run one Newton fit (or, with gd_dims, one block gradient descent) to
    completion on the device; returns (request, final slot).  group:
    workgroups advancing the fit together (None: default_group of the angle
    count; 1: a single workgroup).  The result does not depend on the group
    size."""
    req = FitRequest(atoms, space, data, max_iter, tol, min_iter, seed, record, reg=reg, block=block, perm=perm,
                     h=h, gd_dims=gd_dims)
    dev = req.I.device
    table = torch.as_tensor(np.frombuffer(req.slot_bytes(), dtype=np.uint8).copy(), device=dev)
    budget = budget or max(req.N * req.max_inner * req.nblk, 1)
    group = default_group(req.geom.nA) if group is None else int(group)
    group = max(1, min(group, max_blocks()))
    blocks = torch.as_tensor(block_table([group]), device=dev) if group > 1 else None
    while True:
        if blocks is None:
            run_slots(table, 1, req.dim, budget)
        else:
            run_blocks(table, 1, blocks, group, req.dim, budget)
        s = read_slots(table, 1)[0]
        if s.done or not s.active:
            break
    if s.done == 3:
        raise RuntimeError("the workgroups of the fit failed to meet at a barrier (device oversubscribed?)")
    req.final = s
    return req, s


def _check_fidelity(fidelity):
    if fidelity is not None and not getattr(fidelity, "is_l2_squared", False):
        raise NotImplementedError("the fitters take the l2 fidelity (l2_squared_loss) only")


class Kernels:
    """The three device calls the host-driven fitters make: the sinogram of a
    set of atoms, per-atom L2 derivatives against a sinogram, and per-atom
    gradients of the full loss given its residual."""

    def __init__(self, forward, derivs, resid_grad):
        self.forward, self.derivs, self.resid_grad = forward, derivs, resid_grad


POINT = Kernels(ops.forward, lambda a, s, C, order: ops.derivs(a, s, C, "L2", order),
                lambda a, s, res: ops.derivs(a, s, res, "resid", order=1, want_f=False)[1])
def pixel_kernels(rule="exact"):
    """The point-spread kernels with cell rule `rule` (psf module)."""
    return Kernels(lambda a, s: psf.forward(a, s, rule=rule),
                   lambda a, s, C, order: psf.derivs(a, s, C, "L2", order, rule),
                   lambda a, s, res: psf.derivs(a, s, res, "linear", 1, rule)[1])


PIXEL = pixel_kernels("exact")


def fourier_kernels(vol):
    """The Fourier-plane kernels of a GaussFTVolume: values on its sample
    planes (evaluation cut at |L k|^2 < 30) and the complex L2 derivatives
    (cut at 60), as Radon(.) and Radon.L2_derivs give them in the reference's
    Fourier-domain fits (Fourier_Transform.py:177-207, FTGaussian3D.py:12-92)."""
    from . import fourier
    grid = vol.grid

    def resid_grad(a, s, res):
        # gradient of 1/2 |R - data|^2 for the atoms of a, res = R - data
        return fourier.derivs_all(a, grid, fourier.evaluate(a, grid) - res, 1)[1]

    return Kernels(lambda a, s: fourier.evaluate(a, grid), lambda a, s, C, order: fourier.derivs_all(a, grid, C, order),
                   resid_grad)


def kernels_of(Radon):
    """POINT for GaussTomo (and a bare projection space), the point-spread
    kernels of its rule for SingleParticleGaussTomo, the Fourier-plane kernels
    for GaussFTVolume."""
    from .atomFuncs import SingleParticleGaussTomo
    from .fourier import GaussFTVolume
    if isinstance(Radon, GaussFTVolume):
        return fourier_kernels(Radon)
    if isinstance(Radon, SingleParticleGaussTomo):
        return PIXEL if Radon.rule == "exact" else pixel_kernels(Radon.rule)
    return POINT


def _sumsq(R, data):
    d = R - data
    d = torch.view_as_real(d).double() if d.is_complex() else d.double()
    return 0.5 * float((d * d).sum())


def _device_data(data, dev):
    """data on the device: complex64 for Fourier-plane data, else float32."""
    data = data.data if hasattr(data, "data") and not isinstance(data, torch.Tensor) else data
    data = torch.as_tensor(data, device=dev)
    return data.to(torch.complex64) if data.is_complex() else data.to(torch.float32)


def _project_np(D, iso, th, d):
    """th + d with the projections of GD_lib._step (GD_lib.py:502-522)."""
    t = th + d
    t[0] = max(0.0, t[0])
    r = t[1 + D:]
    if iso:
        r[0] = abs(r[0])
        return t
    sign_sets = ((0, 2), (1,)) if D == 2 else ((0, 3, 5), (1, 4), (2,))
    for s in sign_sets:
        if r[s[0]] < 0:
            r[list(s)] *= -1
    return t


def _project_block_np(D, iso, blk, th, d):
    """Block blk of th = [I, x, r] moved by d with the projections of
    NewtonGaussian.linesearch_block (NewtonGaussian.py:159-187)."""
    t = th.copy()
    o = (0, 1, 1 + D)[blk]
    t[o:o + d.size] += d
    if blk == 0:
        t[0] = max(0.0, t[0])
    elif blk == 1:
        t[1:1 + D] = np.clip(t[1:1 + D], -0.99, 0.99)
    else:
        t = _project_np(D, iso, t, np.zeros_like(t))
        t[0] = th[0]
    return t


def newton_host(recon, space, data, max_iter, kernels=POINT, reg=None, tol=1e-4, min_iter=1, seed=1,
                block=False, rs=None, h=None):
    """NewtonGaussian.linesearch (block=False, NewtonGaussian.py:18-109, with
    GD_lib._step) or linesearch_block (block=True, :112-237) driven from the
    host: per update one forward of the atom, one L2 derivative call against
    data + R_j - R and one forward of the proposal, all on the device; the
    damped system is solved in FP64 on the host.  Same control flow, step
    sizes, projections and energy bookkeeping as the device fitter (fidelity
    in FP64 from the float32 sinograms; E = F + reg; a proposal with a
    non-finite fidelity is rejected).  Fits `recon` in place; returns
    (recon, E, F).  rs (a numpy RandomState) and h ([3 or 1][N] step
    sizes, updated in place) let a caller carry both across calls."""
    reg = as_regulariser(reg)
    dev = recon.I.device
    n = recon.I.numel()
    D = recon.x.shape[1]
    iso = recon.r.reshape(max(n, 1), -1).shape[1] == 1
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    max_iter = list(max_iter) + [1] * (2 - len(max_iter))
    eps = 1e-4
    nblk = 3 if block else 1
    h = np.ones((nblk, n)) if h is None else h
    data = _device_data(data, dev)
    R = kernels.forward(recon, space)
    F = [_sumsq(R, data)]
    E = [F[0] + (reg(recon) if reg.kind else 0.0)]
    rs = np.random.RandomState(seed) if rs is None else rs

    def theta(j):
        return np.concatenate((recon.I[j:j + 1].double().cpu().numpy(), recon.x[j].double().cpu().numpy(),
                               recon.r[j].reshape(-1).double().cpu().numpy()))

    def put(j, th):
        t = torch.as_tensor(th, dtype=torch.float32, device=dev)
        recon.I[j] = t[0]
        recon.x[j] = t[1:1 + D]
        recon.r[j] = t[1 + D:].reshape(recon.r[j].shape)

    for it in range(max_iter[0]):
        brk = [True, True]
        for j in rs.permutation(np.arange(n)):
            one = ops._AtomView(recon.I[j:j + 1], recon.x[j:j + 1], recon.r[j:j + 1])
            for blk in range(nblk):
                for _ in range(max_iter[1]):
                    brk[1] = True
                    Rj = kernels.forward(one, space)
                    _, g, H = kernels.derivs(one, space, data + Rj - R, 2)
                    df, ddf = g[0].cpu().numpy(), H[0].cpu().numpy()
                    if reg.kind:
                        _, gr, Hr = reg.terms(one, 2)
                        df = df + gr[0].cpu().numpy()
                        ddf = ddf + Hr[0].cpu().numpy()
                    ddf[np.diag_indices_from(ddf)] += 1.0 / h[blk, j]
                    old = theta(j)
                    if not block:
                        dx = np.linalg.solve(ddf, -df)
                        new = _project_np(D, iso, old, dx)
                    else:
                        o, m = (0, 1) if blk == 0 else ((1, D) if blk == 1 else (1 + D, old.size - 1 - D))
                        if blk == 0:
                            if not ddf[0, 0] > 1e-8:
                                break
                            d = np.array([-df[0] / ddf[0, 0]])
                        else:
                            try:
                                d = np.linalg.solve(ddf[o:o + m, o:o + m], -df[o:o + m])
                            except np.linalg.LinAlgError:
                                break
                        new = _project_block_np(D, iso, blk, old, d)
                    put(j, new)
                    Rnew = R + kernels.forward(one, space) - Rj
                    Fn = _sumsq(Rnew, data)
                    F.append(Fn)
                    En = Fn + (reg(recon) if reg.kind else 0.0)
                    if En > E[-1] or not np.isfinite(Fn):
                        put(j, old)
                        E.append(E[-1])
                        if h[blk, j] > 2 * eps:
                            brk = [False, False]
                        h[blk, j] /= 10
                    else:
                        R = Rnew
                        E.append(En)
                        if block:
                            # T of NewtonGaussian.py:170-199: I and x unprojected, r projected
                            T = np.float32(old[o:o + m] + d) if blk < 2 else new[o:o + m]
                            moved = np.linalg.norm(old[o:o + m] - T) > tol * np.linalg.norm(old[o:o + m])
                        else:
                            moved = np.sqrt(np.sum(dx * dx)) > tol
                        if moved:
                            brk = [False, False]
                        h[blk, j] *= 1.3 if block else 2
                    h[blk, j] = min(max(h[blk, j], eps), 1 / eps)
                    if brk[1]:
                        break
        if brk[0] and it > min_iter:
            break
    return recon, np.array(E), np.array(F)


def _device_stage(sub, space, data, mi, tol, min_iter, reg, block, rs, h, group):
    """One stage of the greedy path on the device fitter: the permutations
    are drawn from `rs` as the reference's loop would draw them, and `rs` is
    left advanced by the outer iterations the fit actually ran."""
    n = len(sub)
    state = rs.get_state()
    perm = np.stack([rs.permutation(np.arange(n)) for _ in range(mi[0])])
    req, s = newton_fit(sub, space, data, mi, tol, min_iter, reg=reg, block=block, group=group, perm=perm,
                        h=h[:, :n])
    if s.done == 2:
        raise np.linalg.LinAlgError("singular Newton system")
    used = mi[0] if s.it >= mi[0] else s.it + 1
    rs.set_state(state)
    for _ in range(used):
        rs.permutation(np.arange(n))
    h[:, :n] = req.h.reshape(h.shape[0], n).cpu().numpy()
    sub.I.copy_(req.I)
    sub.x.copy_(req.x)
    sub.r.copy_(req.r)
    return req.Ehist[: s.jj + 1].cpu().numpy(), req.Fhist[: s.jj + 1].cpu().numpy()


def greedy(recon, data, max_iter, reg, space, guess, stage):
    """The initial-guess path of NewtonGaussian.linesearch /
    linesearch_block and GD_lib.linesearch_block (NewtonGaussian.py:38-50,
    GD_lib.py:113-126): for n = 1..N, atom n-1 is placed by
    guess(data - Radon(recon[:n]), recon[n-1]), the last energy entry is
    replaced by the energy of recon[:n], and stage(recon[:n], max_iter) runs
    the fitter's sweeps over those atoms (returning its (E, F) histories,
    whose first entry is the stage's start); with three entries in max_iter
    a last stage of max_iter[2] sweeps over all atoms follows.  The stage
    carries the step sizes and the permutation stream from one stage to the
    next, as the reference's single loop does."""
    reg = as_regulariser(reg)
    N = len(recon)
    dev = recon.I.device
    from .spaces import ProjElement
    fwd = stage.kernels.forward
    data = _device_data(data, dev)
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    mi = list(max_iter) + [1] * (2 - len(max_iter))
    E, F = [], []
    n = 1
    while n <= N + 1:
        if n == N + 1:
            if len(mi) == 2:
                break
            n = N
            mi = [mi[2], mi[1]]
        else:
            res = ProjElement(space, data - fwd(recon[:n], space), device=dev)
            recon[n - 1] = guess(res, recon[n - 1])
            sub = recon[:n]
            Fv = _sumsq(fwd(sub, space), data)
            Ev = Fv + (reg(sub) if reg.kind else 0.0)
            if E:
                E[-1], F[-1] = Ev, Fv
            else:
                E, F = [Ev], [Fv]
        sub = recon[:n]
        Es, Fs = stage(sub, mi[:2])
        recon[:n] = sub
        E += list(Es[1:])
        F += list(Fs[1:])
        n += 1
    return recon, np.array(E), np.array(F)


class _Stage:
    """Callable stage for `greedy`, holding the step sizes and the RNG."""

    def __init__(self, kernels, run):
        self.kernels, self.run = kernels, run

    def __call__(self, sub, mi):
        return self.run(sub, mi)


def _check_guess(guess, Radon):
    from .fourier import GaussFTVolume
    if guess is not None and isinstance(Radon, GaussFTVolume):
        raise NotImplementedError("the greedy (guess) path places real-space atoms; it is not defined for "
                                  "Fourier-space fits")


def _newton_dropin(recon, data, max_iter, fidelity, reg, Radon, guess, tol, min_iter, block, kwargs):
    _check_fidelity(fidelity)
    _check_guess(guess, Radon)
    space = Radon.range if Radon is not None else kwargs["space"]
    if guess is not None:
        kern = kernels_of(Radon)
        h = np.ones((3 if block else 1, len(recon)))
        rs = np.random.RandomState(1)
        dev_data = None

        def run(sub, mi):
            nonlocal dev_data
            if kern is POINT:
                if dev_data is None:
                    d = data.data if hasattr(data, "data") and not isinstance(data, torch.Tensor) else data
                    dev_data = torch.as_tensor(d, device=sub.I.device).to(torch.float32)
                return _device_stage(sub, space, dev_data, mi, tol, min_iter, reg, block, rs, h, kwargs.get("group"))
            _, Es, Fs = newton_host(sub, space, data, mi, kern, reg, tol, min_iter, block=block, rs=rs, h=h[:, :len(sub)])
            return Es, Fs

        return greedy(recon, data, max_iter, reg, space, guess, _Stage(kern, run))
    if kernels_of(Radon) is not POINT:
        return newton_host(recon, space, data, max_iter, kernels_of(Radon), reg, tol, min_iter, block=block)
    req, s = newton_fit(recon, space, data, max_iter, tol, min_iter, reg=reg, block=block,
                        group=kwargs.get("group"))
    if s.done == 2:
        raise np.linalg.LinAlgError("singular Newton system")
    recon.I.copy_(req.I)
    recon.x.copy_(req.x)
    recon.r.copy_(req.r)
    E = req.Ehist[: s.jj + 1].cpu().numpy()
    F = req.Fhist[: s.jj + 1].cpu().numpy()
    return recon, E, F


def newton_linesearch(recon, data, max_iter, fidelity=None, reg=None, Radon=None, view=None, guess=None, tol=1e-4,
                      min_iter=1, **kwargs):
    """Drop-in for NewtonGaussian.linesearch: fits `recon` in place and returns
    (recon, E, F): E the energy (fidelity + regulariser) after each update, F
    the fidelity of each update's proposal.  Radon may be GaussTomo (device
    fitter), SingleParticleGaussTomo or GaussFTVolume (host-driven loop over
    that operator's kernels); guess (e.g. a KL initialiser) selects the
    greedy insertion path, max_iter may then have a third entry."""
    return _newton_dropin(recon, data, max_iter, fidelity, reg, Radon, guess, tol, min_iter, False, kwargs)


def newton_linesearch_block(recon, data, max_iter, fidelity=None, reg=None, Radon=None, view=None, guess=None,
                            tol=1e-4, min_iter=1, **kwargs):
    """Drop-in for NewtonGaussian.linesearch_block (NewtonGaussian.py:112-237):
    for each atom the blocks I, x and r in turn, each with its own step size
    (x1.3 on accept, /10 on reject), damped Newton step on its diagonal block
    and projection (I >= 0; x clamped to [-0.99, 0.99]; r as in _step).  The
    reference slices the Hessian with the string 'Ixr' in place of the spatial
    dimension (NewtonGaussian.py:135,163,172), which raises inside its try
    blocks so that its x and r blocks never move; here they do."""
    return _newton_dropin(recon, data, max_iter, fidelity, reg, Radon, guess, tol, min_iter, True, kwargs)


# ---------------------------------------------------------------------------
# Block gradient descent, GD_lib.linesearch_block (GD_lib.py:82-178)
# ---------------------------------------------------------------------------
def gd_linesearch_block(recon, data, max_iter, space, dim="xrI", tol=1e-4, min_iter=1, seed=1, reg=None,
                        return_F=False, kernels=POINT, rs=None, h=None):
    """Block GD over (x, r, I) of each atom with the reference's step-size rules
    (GD_lib.py:96-166: h from 1e-2, h = min(h, 1/max(|g|, 1e-8)), x1.1 on
    accept, /2 on reject, floor 1e-8).  Energies (fidelity + regulariser) in
    FP64.  Divergences from the reference, each a repair:
    * each block steps along its own gradient, T -= h * g_block.  The
      reference selects the gradient by the block's position (df[0],
      df[1:1+D], df[1+D:] for t = 0, 1, 2) while its blocks are ordered
      x, r, I, and moves the whole block by the scalar h * grad[0]
      (GD_lib.py:138-151): for dim='xrI' its x moves along (1, .., 1) dF/dI;
    * on a rejected step the sinogram of the reverted atoms is used again
      (GD_lib.py:157-160 leaves the rejected sinogram in place).
    Returns (recon, E), or (recon, E, F) with F the fidelity of each proposal."""
    reg = as_regulariser(reg)
    dev = recon.I.device
    D = recon.x.shape[1]
    dims = "".join(c for c in "xrI" if c.lower() in dim.lower())
    if not hasattr(max_iter, "__iter__"):
        max_iter = [max_iter, 1]
    max_iter = list(max_iter) + [1] * (2 - len(max_iter))
    eps = 1e-8
    n = recon.I.numel()
    if h is None:
        h = np.full((len(dims), n), 1e-2)
    if kernels is POINT and n > 0:
        return _gd_device(recon, data, max_iter, space, dims, tol, min_iter, seed, reg, return_F, rs, h)
    h = {c: h[t] for t, c in enumerate(dims)}  # rows of the caller's array, updated in place
    data = _device_data(data, dev)
    R = kernels.forward(recon, space)
    F = [_sumsq(R, data)]
    E = [F[0] + (reg(recon) if reg.kind else 0.0)]
    rs = np.random.RandomState(seed) if rs is None else rs
    resid = R - data
    for it in range(max_iter[0]):
        brk = [True, True]
        for j in rs.permutation(np.arange(n)):
            for _ in range(max_iter[1]):
                brk[1] = True
                for name in dims:
                    one = ops._AtomView(recon.I[j:j + 1], recon.x[j:j + 1], recon.r[j:j + 1])
                    g = kernels.resid_grad(one, space, resid)[0].cpu().numpy()
                    if name == "I":
                        grad = g[0:1]
                    elif name == "x":
                        grad = g[1:1 + D]
                    else:
                        grad = g[1 + D:]
                    if reg.kind:  # grad += reg.grad(recon[j], axis=dim[t]) (GD_lib.py:132)
                        grad = grad + np.asarray(reg.grad(one, axis=name)).reshape(-1)
                    arr = {"I": recon.I, "x": recon.x, "r": recon.r}[name]
                    old = arr[j].detach().clone()
                    h[name][j] = min(h[name][j], 1.0 / max(np.sqrt(np.sum(grad * grad)), eps))
                    step = torch.as_tensor(h[name][j] * grad, dtype=torch.float32, device=dev)
                    new = old - (step[0] if name == "I" else step)
                    arr[j] = new
                    Rn = kernels.forward(recon, space)
                    rn = Rn - data
                    Fn = _sumsq(Rn, data)
                    F.append(Fn)
                    En = Fn + (reg(recon) if reg.kind else 0.0)
                    if En > E[-1]:
                        arr[j] = old
                        E.append(E[-1])
                        h[name][j] /= 2
                    else:
                        resid = rn
                        E.append(En)
                        h[name][j] *= 1.1
                        o64, n64 = old.double().cpu().numpy(), new.double().cpu().numpy()
                        if np.sqrt(np.sum((o64 - n64) ** 2)) > tol * np.sqrt(np.sum(o64 * o64)):
                            brk = [False, False]
                    h[name][j] = max(h[name][j], eps)
                if brk[1]:
                    break
        if brk[0] and it > min_iter:
            break
    if return_F:
        return recon, np.array(E), np.array(F)
    return recon, np.array(E)


def _gd_device(recon, data, max_iter, space, dims, tol, min_iter, seed, reg, return_F, rs, h):
    """gd_linesearch_block on the device fitter (fit_kernel, block = 2): the
    atom order of each outer iteration drawn from rs (or RandomState(seed)) as
    the reference's loop draws it, rs left advanced by the outer iterations the
    fit ran and the caller's step sizes h[t][j] updated, as the host loop
    does."""
    n = recon.I.numel()
    rs = np.random.RandomState(seed) if rs is None else rs
    state = rs.get_state()
    perm = np.stack([rs.permutation(np.arange(n)) for _ in range(max_iter[0])])
    dev_data = _device_data(data, recon.I.device)
    req, s = newton_fit(recon, space, dev_data, max_iter, tol, min_iter, reg=reg, perm=perm, h=h, gd_dims=dims)
    used = max_iter[0] if s.it >= max_iter[0] else s.it + 1
    rs.set_state(state)
    for _ in range(used):
        rs.permutation(np.arange(n))
    h[...] = req.h.reshape(len(dims), n).cpu().numpy()
    recon.I.copy_(req.I)
    recon.x.copy_(req.x)
    recon.r.copy_(req.r)
    E = req.Ehist[: s.jj + 1].cpu().numpy()
    if return_F:
        return recon, E, req.Fhist[: s.jj + 1].cpu().numpy()
    return recon, E


def linesearch_block(recon, data, max_iter, fidelity=None, reg=None, Radon=None, view=None, dim="xrI", guess=None,
                     tol=1e-4, min_iter=1, **kwargs):
    """Drop-in for GD_lib.linesearch_block (GD_lib.py:82-178): fits `recon`
    in place, returns (recon, E, F); operators and guess as
    newton_linesearch.  Each block steps along its own gradient (the
    reference takes the gradient by block position and a scalar step,
    GD_lib.py:138-151; see gd_linesearch_block)."""
    _check_fidelity(fidelity)
    _check_guess(guess, Radon)
    space = Radon.range if Radon is not None else kwargs["space"]
    kern = kernels_of(Radon)
    if guess is not None:
        ndim = sum(c.lower() in dim.lower() for c in "xrI")
        h = np.full((ndim, len(recon)), 1e-2)
        rs = np.random.RandomState(1)

        def run(sub, mi):
            _, Es, Fs = gd_linesearch_block(sub, data, mi, space, dim, tol, min_iter, reg=reg, return_F=True,
                                            kernels=kern, rs=rs, h=h[:, :len(sub)])
            return Es, Fs

        return greedy(recon, data, max_iter, reg, space, guess, _Stage(kern, run))
    return gd_linesearch_block(recon, data, max_iter, space, dim, tol, min_iter, reg=reg, return_F=True,
                               kernels=kern)
