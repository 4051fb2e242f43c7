/*
 * gausstomo.h -- C ABI of the MI355X Gaussian-atom tomography library.
 *
 * Every pointer in the structs and arguments below is a DEVICE pointer unless
 * stated otherwise; the library never allocates on these calls.  Work is
 * enqueued on `stream` (a hipStream_t passed as void*, NULL = default stream)
 * and the calls return as soon as it is enqueued.  Return value: 0 on
 * success, GT_EINVAL (-1) for a bad argument, GT_EHIP (-2) for a HIP error,
 * GT_EWORKSPACE (-3) for a workspace smaller than the matching *_workspace()
 * size; gt_last_error() gives the message of the calling thread's last error.
 *
 * The entry points replace the module-level "myRad" functions that
 * GaussTomo._getOps binds (GaussDictCode/atomFuncs.py:106-113):
 *
 *   gt_forward   <- RadProj          GaussDictCode/bin/NewtonBases.py:714-726
 *                                     GaussDictCode/bin/NewtonBases2D.py:274-281
 *   gt_derivs    <- L2derivs_RadProj GaussDictCode/bin/NewtonBases.py:759-789
 *                                     GaussDictCode/bin/NewtonBases2D.py:299-326
 *                <- derivs_RadProj   GaussDictCode/bin/NewtonBases.py:954-990
 *                                     GaussDictCode/bin/NewtonBases2D.py:415-441
 *                   (batched over atoms: GaussTomo.derivs / L2_derivs loop,
 *                    GaussDictCode/atomFuncs.py:130-158)
 *   gt_volume    <- VolProj          GaussDictCode/bin/NewtonBases.py:691-695
 *                                     GaussDictCode/bin/NewtonBases2D.py:252-256
 *   gt_ft_eval   <- evaluate         GaussDictCode/bin/FTGaussian3D.py:12-52 (grid)
 *   gt_ft_planes_eval <- evaluate    GaussDictCode/bin/FTGaussian3D.py:12-38 (planes),
 *                                     GaussDictCode/bin/FTGaussian2D.py:13-35
 *   gt_ft_derivs <- derivs           GaussDictCode/bin/FTGaussian3D.py:55-92
 */
#ifndef GAUSSTOMO_H
#define GAUSSTOMO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GT_EINVAL (-1)
#define GT_EHIP (-2)
#define GT_EWORKSPACE (-3)

/* Derivative modes of gt_derivs. */
#define GT_MODE_LINEAR 0 /* f = sum R C, Df = sum C dR, DDf = sum C ddR       (derivs_RadProj)   */
#define GT_MODE_L2 1     /* f = 1/2 sum (R-C)^2, Df, DDf of that              (L2derivs_RadProj) */
#define GT_MODE_RESID 2  /* C is the full residual R_all - data: Df, DDf of    */
                         /* 1/2 ||R_all - data||^2 w.r.t. each atom (= L2 with */
                         /* C_j = data - R_all + R_j, NewtonGaussian.py:59-60) */

/* Atoms in the reference layout (GaussDictCode/dictionary_def.py:147-214). */
typedef struct gt_atoms {
    int dim;        /* 2 or 3 */
    int n;          /* number of atoms (>= 0) */
    int iso;        /* 1: r holds one radius per atom, 0: dim*(dim+1)/2 slots */
    const float *I; /* [n] */
    const float *x; /* [n][dim] */
    const float *r; /* [n][iso ? 1 : dim*(dim+1)/2], slots [r00,r11,(r22,)r01,(r12,r02)] */
} gt_atoms;

/* Projection geometry (ProjSpace, GaussDictCode/dictionary_def.py:263-283). */
typedef struct gt_geom {
    int dim;         /* 2 or 3 */
    int lenT;        /* 3: 3D (w0, w1 given); 2: 2D, or 3D single-axis ("2.5D") */
    int nA;          /* number of orientations */
    int n0, n1;      /* detector sizes; n1 = 1 when dim == 2 */
    int sorted;      /* 1 if p0 (and p1) are ascending: enables footprint culling */
    const float *T;  /* [nA][lenT] ray directions (orientations) */
    const float *w0; /* [nA][lenT] first detector axis (ortho[0]) */
    const float *w1; /* [nA][3] second detector axis (ortho[1]); 3D only, else NULL */
    const float *p0; /* [n0] detector coordinates (detector[0]) */
    const float *p1; /* [n1] detector coordinates (detector[1]); dim 3 only */
} gt_geom;

/* Volume grid for gt_volume (VolSpace coordinate vectors). */
typedef struct gt_vol {
    int dim;
    int m0, m1, m2;  /* m2 = 1 when dim == 2 */
    int sorted;      /* 1 if the coordinate vectors are ascending */
    const float *y0, *y1, *y2;
} gt_vol;

/* Separable Fourier grid for gt_ft_eval: k = (k0[i0], k1[i1], k2[i2]). */
typedef struct gt_kgrid {
    int dim; /* 3 */
    int m0, m1, m2;
    int sorted;
    const float *k0, *k1, *k2;
} gt_kgrid;

const char *gt_last_error(void);
const char *gt_version(void);

/* Kernel ids for gt_profile_read. */
#define GT_K_SETUP 0   /* per-(atom, angle) FP64 setup            */
#define GT_K_FORWARD 1 /* forward projection                      */
#define GT_K_MOMENTS 2 /* derivative pixel moments                */
#define GT_K_FINISH 3  /* FP64 moment -> gradient / Hessian       */
#define GT_K_VOLUME 4  /* volume render                           */
#define GT_K_FT 5      /* Fourier evaluation                      */
#define GT_K_FIT 6     /* fitter kernels                          */

/* Timing support (no counterpart in the reference, which only prints wall
 * time): when enabled, each kernel launch is bracketed by HIP events on its
 * stream; gt_profile_read waits for the recorded launches of one kernel id,
 * returns their summed duration in ms and their count, and forgets them.
 * gt_profile_select limits the bracketing to the kernel ids whose bit
 * (1u << id) is set in mask (all of them by default): each event pair adds
 * a few microseconds between launches, so a timed run brackets only the
 * kernel it measures. */
int gt_profile_enable(int on);
int gt_profile_select(unsigned mask);
int gt_profile_read(int kernel_id, double *total_ms, long *count);

/* RadProj: R[a][i0][i1] = sum over atoms of the line integral (written, not accumulated). */
size_t gt_forward_workspace(const gt_atoms *atoms, const gt_geom *geom);
int gt_forward(const gt_atoms *atoms, const gt_geom *geom, float *R, void *ws, size_t ws_bytes,
               void *stream);

/* Batched per-atom derivatives over all rays.  C: [nA][n0][n1] FP32.
 * Outputs (FP64, any may be NULL except g): f[n], g[n][P], H[n][P][P] with
 * P = 1 + dim + (iso ? 1 : dim*(dim+1)/2), parameter order [I, x, r].
 * order 1 skips H. */
size_t gt_derivs_workspace(const gt_atoms *atoms, const gt_geom *geom, int mode);
int gt_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order, double *f,
              double *g, double *H, void *ws, size_t ws_bytes, void *stream);

/* One sweep of the l2 fit in one call: R = RadProj(atoms) - data, then for
 * every atom the gradient g[n][P] and (order 2) Hessian block H[n][P][P] of
 * 1/2 ||R||^2 with respect to that atom (what NewtonGaussian.linesearch and
 * GD_lib.linesearch_block request atom by atom through L2derivs_RadProj with
 * C = data - R + R_j, NewtonGaussian.py:59-60, GD_lib.py:121-123), and, if f is
 * not NULL, f[0] = 1/2 ||R||^2 in FP64.  The records are set up once for the
 * projection and the derivatives.  Workspace: gt_derivs_workspace(.., GT_MODE_RESID).
 * R and data are [nA][n0][n1] device arrays and must not overlap. */
int gt_forward_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *data, int order, float *R,
                      double *f, double *g, double *H, void *ws, size_t ws_bytes, void *stream);

/* No counterpart in the reference: gt_forward_derivs without a workspace, for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The workspace is a
 * file-scope device array of GT_NOALLOC_SCRATCH_BYTES that each GPU the
 * library runs on holds until the process ends; the call allocates nothing
 * and clears nothing.  A sweep whose workspace (gt_derivs_workspace(..,
 * GT_MODE_RESID)) exceeds it returns GT_EWORKSPACE.  Calls that overlap on
 * one GPU (other streams, other host threads) race on the array. */
#define GT_NOALLOC_SCRATCH_BYTES ((size_t)128 << 20)
size_t gt_noalloc_scratch_bytes(void);
int gt_forward_derivs_noalloc(const gt_atoms *atoms, const gt_geom *geom, const float *data, int order, float *R,
                              double *f, double *g, double *H, void *stream);

/* VolProj: u[j0][j1][j2] = sum over atoms of I sqrt(30/pi) exp(-30|U(y-x)|^2), cut at 1. */
size_t gt_volume_workspace(const gt_atoms *atoms, const gt_vol *vol);
int gt_volume(const gt_atoms *atoms, const gt_vol *vol, float *u, void *ws, size_t ws_bytes, void *stream);

/* Fourier evaluation on a separable 3D grid (FTGaussian3D.__grid_eval_*,
 * GaussDictCode/bin/FTGaussian3D.py:95-110,151-164).  atoms hold FT-space
 * parameters (Fourier_Transform.py:293-333).  out: [m0][m1][m2] complex64
 * as interleaved (re, im) floats. */
size_t gt_ft_workspace(const gt_atoms *atoms, const gt_kgrid *grid);
int gt_ft_eval(const gt_atoms *atoms, const gt_kgrid *grid, float *out, void *ws, size_t ws_bytes,
               void *stream);

/*
 * Point-spread ("single particle") projector, the operator
 * SingleParticleGaussTomo binds (GaussDictCode/atomFuncs.py:161-172):
 *   gt_psf_forward  <- RadProj          GaussDictCode/bin/point_spread_functions.py:371-385
 *   gt_psf_derivs   <- L2derivs_RadProj :1077-1115 (mode GT_MODE_L2)
 *                   <- derivs_RadProj   :1314-1350 (mode GT_MODE_LINEAR)
 *   gt_psf_volume   <- VolProj          :168-197
 * Each detector value is the integral of the atom's line integral over its
 * pixel (the square of side |p[1] - p[0]| around the detector coordinate,
 * per axis); each volume value the integral of the atom's profile over its
 * voxel.  Detector and volume coordinates must ascend with at least two
 * entries per axis.  Outputs and workspaces as gt_forward / gt_derivs /
 * gt_volume (derivatives batched over atoms, FP64, iso collapsed).
 */
size_t gt_psf_forward_workspace(const gt_atoms *atoms, const gt_geom *geom);
int gt_psf_forward(const gt_atoms *atoms, const gt_geom *geom, float *R, void *ws, size_t ws_bytes, void *stream);
size_t gt_psf_derivs_workspace(const gt_atoms *atoms, const gt_geom *geom);
int gt_psf_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order, double *f,
                  double *g, double *H, void *ws, size_t ws_bytes, void *stream);
size_t gt_psf_volume_workspace(const gt_atoms *atoms, const gt_vol *vol);
int gt_psf_volume(const gt_atoms *atoms, const gt_vol *vol, float *u, void *ws, size_t ws_bytes, void *stream);

/*
 * The same three operators with a choice of cell rule (workspaces as above):
 *   GT_PSF_EXACT    the pixel / voxel integral (what gt_psf_forward,
 *                   gt_psf_derivs and gt_psf_volume compute);
 *   GT_PSF_SAMPLED  the reference's own numbers: __rf2 / __rf3
 *                   (point_spread_functions.py:413-527) for RadProj,
 *                   __drf2 / __drf3 (:924-1075) for the derivatives and
 *                   __if2 / __if3 (:222-291) for VolProj.  Per (atom, cell),
 *                   with n the distance of the (projected) atom centre from
 *                   the cell centre and R = 1 / min diag U: n + R < s gives
 *                   the atom's total I (pi/30) / det U (2D: I sqrt(pi/30) /
 *                   det U) and no derivative, n - R > s gives 0, otherwise
 *                   the integrand is sampled at lo + k (hi - lo) / 10,
 *                   k = 1..11, per axis of [lo, hi] = the cell cut to
 *                   [-R, R] about the centre, each sample weighted
 *                   (hi - lo) / 11; s is half the first detector axis'
 *                   spacing on both detector axes, half the second volume
 *                   axis' spacing on every volume axis.  Detector frames are
 *                   taken as orthonormal (n from the detector offsets).
 *                   Isotropic derivatives use the exact chain rule (the
 *                   reference's collapse, :1101-1113, drops terms).
 * Any other rule value returns GT_EINVAL.
 */
#define GT_PSF_EXACT 0
#define GT_PSF_SAMPLED 1
int gt_psf_forward_rule(const gt_atoms *atoms, const gt_geom *geom, int rule, float *R, void *ws, size_t ws_bytes,
                        void *stream);
int gt_psf_derivs_rule(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order, int rule,
                       double *f, double *g, double *H, void *ws, size_t ws_bytes, void *stream);
int gt_psf_volume_rule(const gt_atoms *atoms, const gt_vol *vol, int rule, float *u, void *ws, size_t ws_bytes,
                       void *stream);

/*
 * Sinogram moments for the KL moment initialiser (KL_GaussRadon.py:57-190):
 * replaces MOM0 / MOM1 / MOM2 (:221-392) and the f updates of
 * doKL_ProjGDStep_iso (:128) and doKL_LagrangeStep_iso (Newton_sino_2D,
 * :193-229).  Per angle a, FP64, into out[a][5] (device):
 *   sum f, sum f p0, sum f p1, sum f p0^2, sum f p1^2   (p1 terms 0 in 2D)
 * over the detector, with f formed per pixel from S and Q ([nA][n0][n1],
 * device, float32):
 *   GT_KL_CLAMP     f = max(0, a S + b Q)        (Q may be NULL: b ignored)
 *   GT_KL_LAGRANGE  f + a log f = S + a log(b Q), a >= 0 (f = S when a == 0,
 *                   f = 0 where b Q <= 1e-15)
 * Only geom->dim, nA, n0, n1, p0 and p1 are used.  Deterministic (fixed-order
 * reduction).
 */
#define GT_KL_CLAMP 0
#define GT_KL_LAGRANGE 1
size_t gt_sino_moments_workspace(const gt_geom *geom);
int gt_sino_moments(const gt_geom *geom, const float *S, const float *Q, int mode, double a, double b, double *out,
                    void *ws, size_t ws_bytes, void *stream);

/*
 * Fourier evaluation on sample planes (FTGaussian3D.evaluate hyperplane and
 * 2.5D branches, GaussDictCode/bin/FTGaussian3D.py:12-38,113-148,167-196;
 * FTGaussian2D.evaluate, GaussDictCode/bin/FTGaussian2D.py:13-35,53-100):
 *   out[a][i0][i1] = sum over FT-space atoms of I exp(-|L k|^2/2) exp(-i x.k),
 *   |L k|^2 < 30 (3D atoms) or 60 (2D atoms), rk = L k with L lower-triangular
 *   from the slots (3D: rk_1 = r3 k_0 + r1 k_1, rk_2 = r5 k_0 + r4 k_1 + r2 k_2;
 *   2D: rk_1 = r2 k_0 + r1 k_1); isotropic: |k|^2 / r^2 (3D), r^2 |k|^2 (2D).
 * Sample forms (k for sample (a, i0, i1)):
 */
#define GT_KP_3D 0      /* dim 3: k0[i0] w0[a] + k1[i1] w1[a]; w0, w1 [nA][3]               */
#define GT_KP_2P5D 1    /* dim 3: (k0[i0] w0[a][0], k0[i0] w0[a][1], k1[i1]); w0 [nA][2]     */
#define GT_KP_2D_LINE 2 /* dim 2: k0[i0] w0[a]; w0 [nA][2]; m1 = 1, k1 unused                */
#define GT_KP_2D_GRID 3 /* dim 2: (k0[i0], k1[i1]); nA = 1, w0 / w1 unused                   */
typedef struct gt_kplanes {
    int form;
    int nA, m0, m1;
    int sorted; /* 1 if k0 (and k1) ascend: enables exact culling */
    const float *k0, *k1, *w0, *w1;
} gt_kplanes;
/* FP64 variants (the reference's 2D evaluator has float64 signatures). */
typedef struct gt_kplanes64 {
    int form;
    int nA, m0, m1;
    int sorted;
    const double *k0, *k1, *w0, *w1;
} gt_kplanes64;
typedef struct gt_atoms64 {
    int dim, n, iso;
    const double *I, *x, *r;
} gt_atoms64;

/* out: [nA][m0][m1] complex64 (interleaved floats) / complex128. */
size_t gt_ft_planes_workspace(int n_atoms, const gt_kplanes *planes);
int gt_ft_planes_eval(const gt_atoms *atoms, const gt_kplanes *planes, float *out, void *ws, size_t ws_bytes,
                      void *stream);
size_t gt_ft_planes_workspace_f64(int n_atoms, const gt_kplanes64 *planes);
int gt_ft_planes_eval_f64(const gt_atoms64 *atoms, const gt_kplanes64 *planes, double *out, void *ws,
                          size_t ws_bytes, void *stream);

/* Complex L2 derivatives over hyperplane samples (FTGaussian3D.derivs,
 * GaussDictCode/bin/FTGaussian3D.py:55-92,199-652), for every 3D FT-space atom
 * independently against all of C ([nA][m0][m1] complex64, forms GT_KP_3D /
 * GT_KP_2P5D): f[n] = 1/2 sum |g - C|^2, g[n][P] = Re sum (g - C) conj(dg),
 * H[n][P][P] = Re sum dg conj(dg)^T + (g - C) conj(ddg), g cut where
 * |L k|^2 > 60; P = 10 (iso 5), order [I, x, slots].  FP64 outputs; f and H may
 * be NULL.  Workspace: gt_ft_derivs_workspace. */
size_t gt_ft_derivs_workspace(const gt_atoms *atoms, const gt_kplanes *planes);
int gt_ft_derivs(const gt_atoms *atoms, const gt_kplanes *planes, const float *C, double *f, double *g, double *H,
                 void *ws, size_t ws_bytes, void *stream);

/*
 * Regularisers (GaussDictCode/regularisation.py:22-442).  Each is a sum over
 * atoms of one term per atom; params (GT_REG_NPARAM doubles, unused ones 0):
 *   GT_REG_NULL     -                                    null      :22-38
 *   GT_REG_JOUBERT  a, b, c, d, e: -a log I + b/2|x|^2 + c v - d log v - e,
 *                   v = r^dim (iso) or the product of the diagonal slots;
 *                   from Joubert(dim, intensity, location, (r0, r1)):
 *                   a = intensity, b = location, c = r1, d = r1 r0,
 *                   e = r0 r1 (1 - log r0)                Joubert   :41-108
 *   GT_REG_MASS     m, k, s: s (|I R| - m)^k, R = r^dim (iso) or
 *                   1 / product of the diagonal slots     Mass      :111-210
 *   GT_REG_VOLUME   r0^dim, k, s: s |v - r0^dim|^k, v = r^-dim (iso) or the
 *                   product of the diagonal slots         Volume    :213-318
 *   GT_REG_RADIUS   r0, s, -k: s sum (r_i - r0)^-k over the diagonal slots
 *                   (iso: dim (r - r0)^-k)                Radius    :321-386
 *   GT_REG_ISO      s: s/2 sum of the squared off-diagonal slots  Iso :389-442
 * A term is +inf outside its domain (Joubert: I <= 0 or a diagonal slot <= 0;
 * Mass: I < 0; Volume: a diagonal slot <= 0; Radius: a diagonal slot <= r0).
 * Gradients and Hessians are the exact derivatives of the term.
 */
#define GT_REG_NULL 0
#define GT_REG_JOUBERT 1
#define GT_REG_MASS 2
#define GT_REG_VOLUME 3
#define GT_REG_RADIUS 4
#define GT_REG_ISO 5
#define GT_REG_NPARAM 5

/* Per-atom regulariser terms of all atoms (reg(atoms), reg.grad, reg.hess,
 * regularisation.py:22-442, for one atom at a time there), FP64:
 * val[n], and unless NULL grad[n][P] and hess[n][P][P], P = 1 + dim +
 * (iso ? 1 : dim*(dim+1)/2).  params is a HOST array of GT_REG_NPARAM. */
int gt_reg_eval(const gt_atoms *atoms, int kind, const double *params, double *val, double *grad, double *hess,
                void *stream);

/*
 * Device-resident damped-Newton fitter (NewtonGaussian.linesearch with
 * GD_lib._step, NewtonGaussian.py:18-109, GD_lib.py:492-536, or with block =
 * 1 NewtonGaussian.linesearch_block, :112-237; any of the regularisers
 * above, l2 fidelity).  One gt_fit_slot per fit request, in device
 * memory; one workgroup (gt_fit_run) or a group of workgroups
 * (gt_fit_run_blocks) advances one slot.  The host fills the problem and
 * state pointers, the initial residual (forward - data) and energy, and the
 * per-outer-iteration atom permutations; gt_fit_run then performs up to
 * `budget` atom updates on every active slot and returns (stream-ordered),
 * leaving the progress fields updated.  Slots with done != 0 or active == 0
 * are skipped, so the same launch (or a captured graph of it) can be replayed
 * while requests are admitted into and retired from the table.
 */
typedef struct gt_fit_slot {
    /* problem (read-only for the kernel) */
    int active;       /* 0: slot unused */
    int iso;          /* atoms.iso */
    int N, nA, n0, n1;
    int lenT, sorted;
    int max_outer, max_inner, min_iter;
    int quota;        /* if > 0: at most this many atom updates per gt_fit_run (else `budget`) */
    double tol;       /* convergence threshold: on ||dx|| (block = 0), on
                         ||new - old|| / ||old|| of the updated block (block = 1) */
    const float *T, *w0, *w1, *p0, *p1; /* geometry (gt_geom layout) */
    const int *perm;  /* [max_outer][N] atom order of each outer iteration */
    /* state (read-write) */
    float *I, *x, *r; /* atoms, gt_atoms layout */
    float *resid;     /* [nA][n0][n1] current projection minus data */
    double *h;        /* [block ? 3 : 1][N] per-atom step sizes, initialised to 1 */
    double *Ehist;    /* [(block ? 3 : 1)*max_outer*max_inner*N + 1] energy after each update, or NULL */
    float *scratch;   /* gt_fit_scratch_bytes(nA) bytes, zeroed before the first run */
    float *stage;     /* [nA][n0][n1] old residual of the update in flight, or NULL: with it
                         the energy pass commits the proposal at once and a rejected one is
                         restored from here (same results, one pixel pass less) */
    /* progress (read-write) */
    int it, pos, inner, jj;  /* outer iteration, position in perm, inner step, history index */
    int brk0, brk1;          /* convergence flags of the current outer / inner loop */
    int done;                /* 1: converged or out of iterations; 2: singular solve (block = 0);
                                3: the workgroups of a group failed to meet (see gt_fit_run_blocks) */
    int pad1;
    double E;                /* current energy: fidelity + regulariser */
    long long updates;       /* atom updates performed */
    long long accepted;      /* accepted updates */
    long long *phase_ticks;  /* [8] optional: wall-clock ticks (100 MHz) spent per phase, or NULL */
    /* variant and regulariser (read-only) */
    int block;               /* 0: one Newton step in all parameters per update (linesearch);
                                1: I, then x, then r of the atom, each with its own step size,
                                projection and inner loop (linesearch_block);
                                2: block gradient descent (GD_lib.linesearch_block): the
                                gd_nblk blocks gd_blocks[] in turn inside each inner iteration,
                                h[gd_nblk][N] from 1e-2 */
    int reg_kind;            /* GT_REG_* */
    double reg_p[GT_REG_NPARAM];
    double *Fhist;           /* like Ehist: fidelity of each update's proposal (also when
                                rejected, as the reference's F records), or NULL */
    /* regularised energy (read-write): E = F + (ninf ? +inf : Rsum) */
    int blk;                 /* current parameter block (block variant), 0 otherwise */
    int ninf;                /* atoms whose regulariser term is not finite */
    double F;                /* fidelity 1/2 ||resid||^2 */
    double Rsum;             /* sum of the finite regulariser terms */
    /* block gradient descent (block = 2, read-only): the parameter blocks in
       the order of the reference's dims string, 0 = I, 1 = x, 2 = r */
    int gd_nblk;
    int gd_blocks[3];
} gt_fit_slot;

/* nA must be <= GT_FIT_MAX_ANGLES and n0, n1 <= GT_FIT_MAX_DETECTOR. dim 2 or 3. */
#define GT_FIT_MAX_ANGLES 256
#define GT_FIT_MAX_DETECTOR 1024
size_t gt_fit_scratch_bytes(int nA);
int gt_fit_run(gt_fit_slot *slots, int nslots, int dim, int budget, void *stream);

/*
 * Grouped launch: workgroup b of the grid advances slot blocks[b].slot as
 * member blocks[b].member of a group of blocks[b].group workgroups, which
 * share the slot's angles (member m owns angles m, m + group, ...) and meet
 * at two device-wide barriers per atom update.  Gradients, Hessians and
 * energy changes are reduced per angle in a fixed order, so a fit's result
 * does not depend on its group size.  Every member of a slot must appear
 * exactly once with the same group and quota (quota > 0: atom updates this
 * launch, else the slot's quota / budget).  slot < 0 marks an idle block.
 * nblocks must not exceed the number of workgroups the device holds at once
 * (one per CU); a member that waits more than about two seconds for the
 * others marks the slot done = 3 and returns.  gt_fit_run is the same
 * launch with one workgroup per slot.  blocks is a device array.
 */
typedef struct gt_fit_block {
    int slot, member, group, quota;
} gt_fit_block;
int gt_fit_max_blocks(void);
int gt_fit_run_blocks(gt_fit_slot *slots, int nslots, const gt_fit_block *blocks, int nblocks, int dim, int budget,
                      void *stream);

#ifdef __cplusplus
}
#endif
#endif
