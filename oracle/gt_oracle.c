/*
 * gt_oracle.c -- FP64 CPU restatement of the reference's Gaussian-atom
 * projector kernels.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py may load it.  The product path never calls it.
 *
 * It follows the reference's dense loops literally (one ray x all atoms for
 * the forward, one atom x all rays for the derivatives), in double precision:
 *
 *   Rg / g profiles          GaussDictCode/bin/NewtonBases.py:40-71
 *   __f                      GaussDictCode/bin/NewtonBases.py:237-257
 *                            GaussDictCode/bin/NewtonBases2D.py:16-34
 *   __Rf                     GaussDictCode/bin/NewtonBases.py:260-310
 *                            GaussDictCode/bin/NewtonBases2D.py:38-72
 *   __dRf_aniso / __dRf      GaussDictCode/bin/NewtonBases.py:313-527
 *                            GaussDictCode/bin/NewtonBases2D.py:76-242
 *   __tovec3D / __tovec2p5D  GaussDictCode/bin/NewtonBases.py:676-688
 *   __tovec                  GaussDictCode/bin/NewtonBases2D.py:246-249
 *   __RadProj_3D / _2p5D     GaussDictCode/bin/NewtonBases.py:729-756
 *   __RadProj (2D)           GaussDictCode/bin/NewtonBases2D.py:284-296
 *   __L2derivs_RadProj_*     GaussDictCode/bin/NewtonBases.py:792-870
 *                            GaussDictCode/bin/NewtonBases2D.py:329-401
 *   __derivs_RadProj_*       GaussDictCode/bin/NewtonBases.py:993-1069
 *                            GaussDictCode/bin/NewtonBases2D.py:444-515
 *   __VolProj                GaussDictCode/bin/NewtonBases.py:698-711
 *                            GaussDictCode/bin/NewtonBases2D.py:259-271
 *   __grid_eval_* / __hyperplane_eval_*
 *                            GaussDictCode/bin/FTGaussian3D.py:95-196
 *   __*_FTGaussEval_*_CPU    GaussDictCode/bin/FTGaussian2D.py:42-100
 *
 * The derivative kernels are written once for dimension D in {2,3} using the
 * upper-triangular matrix U of the atom, which is what the reference's
 * unrolled index formulas expand to (the x-x block is
 * I * sum_kl ddJdyy[k][l] U[k][a] U[l][b], the x-r block
 * -I * sum_l (ddJdyy[i0][l] Y[i1] + ddJdyT[l][i0] T[i1]) U[l][a] - [a==i1] I dJdy[i0],
 * and so on).  Isotropic atoms are handled by the Python side as U = rho*Id
 * plus the chain rule, which is the intended mathematics of the (broken)
 * reference iso kernels.
 *
 * Rays whose quadratic m exceeds the cutoff contribute exactly zero to R, dR
 * and ddR in the reference (the dg == 0 branch, NewtonBases.py:375-380); the
 * loops below skip the derivative algebra for them, which changes nothing.
 */
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>
#undef I /* complex.h's imaginary unit; I is the atom intensity here */

/* Arithmetic type: double for the checker.  The CPU baseline also builds this
 * file with -DGT_REAL=float -fsingle-precision-constant, the reference's own
 * precision (float32 kernels, NewtonBases.py:729-1069). */
#ifndef GT_REAL
#define GT_REAL double
#endif
typedef GT_REAL real;

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* r-slot -> (row, col) of U.  NewtonBases.py:319-325, NewtonBases2D.py:84-87 */
static const int RIDX3[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {1, 2}, {0, 2}};
static const int RIDX2[3][2] = {{0, 0}, {1, 1}, {0, 1}};

static void fill_U(int D, const real *r, real U[3][3])
{
    memset(U, 0, sizeof(real) * 9);
    int nr = D * (D + 1) / 2;
    for (int k = 0; k < nr; k++) {
        int a = (D == 3) ? RIDX3[k][0] : RIDX2[k][0];
        int b = (D == 3) ? RIDX3[k][1] : RIDX2[k][1];
        U[a][b] = r[k];
    }
}

/* __g: NewtonBases.py:40-46 */
static inline real prof_g(real n) { return n > 1.0 ? 0.0 : exp(-30.0 * n) * sqrt(30.0 / M_PI); }

/* __f: NewtonBases.py:237-257 (D=3), NewtonBases2D.py:16-34 (D=2) */
static real atom_f(int D, real I, const real *x, const real U[3][3], const real *y)
{
    real Y[3], n = 0.0;
    for (int i = 0; i < D; i++) Y[i] = y[i] - x[i];
    for (int i = 0; i < D; i++) {
        real s = 0.0;
        for (int j = i; j < D; j++) s += U[i][j] * Y[j];
        n += s * s;
    }
    return I * prof_g(n);
}

/* Shared front half of __Rf and __dRf: Y, rY = U Y, rT = U T / |U T|, n = 1/|U T|,
 * IP = rT.rY, MY = rY - rT IP, m = |MY|^2.  NewtonBases.py:267-308. */
struct ray_geom {
    real Y[3], T[3], rY[3], rT[3], MY[3], n, IP, m;
};

static void ray_setup(int D, const real *x, const real U[3][3], const real *y,
                      const real *T, struct ray_geom *g)
{
    for (int i = 0; i < 3; i++) {
        g->Y[i] = (i < D) ? y[i] - x[i] : 0.0;
        g->T[i] = (i < D) ? T[i] : 0.0;
    }
    real nn = 0.0;
    for (int i = 0; i < D; i++) {
        real sy = 0.0, st = 0.0;
        for (int j = i; j < D; j++) {
            sy += U[i][j] * g->Y[j];
            st += U[i][j] * g->T[j];
        }
        g->rY[i] = sy;
        g->rT[i] = st;
        nn += st * st;
    }
    g->n = 1.0 / sqrt(nn);
    g->IP = 0.0;
    for (int i = 0; i < D; i++) {
        g->rT[i] *= g->n;
        g->IP += g->rT[i] * g->rY[i];
    }
    g->m = 0.0;
    for (int i = 0; i < D; i++) {
        g->MY[i] = g->rY[i] - g->rT[i] * g->IP;
        g->m += g->MY[i] * g->MY[i];
    }
}

/* __Rf: I * Rg(m) * n.  NewtonBases.py:260-310, NewtonBases2D.py:38-72 */
static real atom_Rf(int D, real I, const real *x, const real U[3][3], const real *y,
                      const real *T)
{
    struct ray_geom g;
    ray_setup(D, x, U, y, T, &g);
    return g.m > 1.0 ? 0.0 : I * exp(-30.0 * g.m) * g.n;
}

/* __dRf_aniso / __dRf: value, gradient and Hessian of R w.r.t. [I, x, r].
 * NewtonBases.py:313-527, NewtonBases2D.py:76-242.  Returns 0 (and leaves the
 * outputs untouched) if the ray is outside the cutoff. */
static int atom_dRf(int D, real I, const real *x, const real *r, const real U[3][3],
                    const real *y, const real *T, int order, real *R, real *dR,
                    real *ddR /* P x P row-major */)
{
    const int nr = D * (D + 1) / 2, P = 1 + D + nr;
    struct ray_geom g;
    ray_setup(D, x, U, y, T, &g);
    if (g.m > 1.0) return 0;
    (void)r;

    const real *MY = g.MY, *rT = g.rT, *Y = g.Y, *Tv = g.T;
    const real n = g.n, IP = g.IP;
    real dg0 = exp(-30.0 * g.m), dg1 = -30.0 * dg0, dg2 = 900.0 * dg0; /* __dRg :59-71 */

    real dMydT[3][3], dJdy[3], dJdT[3];
    real ddJdyy[3][3], ddJdyT[3][3], ddJdTT[3][3];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++)
            dMydT[i][j] = rT[i] * (IP * rT[j] - MY[j]) - (i == j ? IP : 0.0);
    for (int i = 0; i < D; i++) {
        dJdy[i] = 2.0 * n * dg1 * MY[i];
        dJdT[i] = -n * n * (2.0 * IP * MY[i] * dg1 + rT[i] * dg0);
    }
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            real d = (i == j) ? 1.0 : 0.0;
            ddJdyy[i][j] = 2.0 * n * (2.0 * MY[i] * MY[j] * dg2 + (d - rT[i] * rT[j]) * dg1);
            ddJdyT[i][j] = 2.0 * n * n * ((dMydT[i][j] - MY[i] * rT[j]) * dg1 -
                                          2.0 * IP * MY[i] * MY[j] * dg2);
        }
    /* ddJdTT: the reference evaluates its (i<=j) formula and mirrors it. */
    for (int i = 0; i < D; i++)
        for (int j = i; j < D; j++) {
            real d = (i == j) ? 1.0 : 0.0;
            real v = 2.0 * n * n * n *
                       ((1.5 * rT[i] * rT[j] - 0.5 * d) * dg0 +
                        (2.0 * IP * MY[i] * rT[j] + IP * rT[i] * MY[j] - IP * dMydT[i][j] -
                         MY[i] * MY[j]) * dg1 +
                        2.0 * IP * IP * MY[i] * MY[j] * dg2);
            ddJdTT[i][j] = v;
            ddJdTT[j][i] = v;
        }

    const int (*RIDX)[2] = (D == 3) ? RIDX3 : RIDX2;
    memset(dR, 0, sizeof(real) * P);
    memset(ddR, 0, sizeof(real) * P * P);

    *R = I * dg0 * n;
    dR[0] = dg0 * n;
    /* d/dI d/dx: -sum_l U[l][a] dJdy[l] */
    for (int a = 0; a < D; a++) {
        real s = 0.0;
        for (int l = 0; l < D; l++) s += U[l][a] * dJdy[l];
        ddR[0 * P + 1 + a] = -s;
    }
    /* d/dI d/dr */
    for (int k = 0; k < nr; k++) {
        int i0 = RIDX[k][0], i1 = RIDX[k][1];
        ddR[0 * P + 1 + D + k] = dJdy[i0] * Y[i1] + dJdT[i0] * Tv[i1];
    }
    for (int i = 1; i < P; i++) dR[i] = I * ddR[0 * P + i];

    if (order > 1) {
        /* r-r block */
        for (int k = 0; k < nr; k++) {
            int i0 = RIDX[k][0], i1 = RIDX[k][1];
            for (int l = k; l < nr; l++) {
                int j0 = RIDX[l][0], j1 = RIDX[l][1];
                ddR[(1 + D + k) * P + 1 + D + l] =
                    I * (ddJdyy[i0][j0] * Y[i1] * Y[j1] + ddJdyT[i0][j0] * Y[i1] * Tv[j1] +
                         ddJdyT[j0][i0] * Y[j1] * Tv[i1] + ddJdTT[i0][j0] * Tv[i1] * Tv[j1]);
            }
        }
        /* x-r block */
        for (int k = 0; k < nr; k++) {
            int i0 = RIDX[k][0], i1 = RIDX[k][1];
            for (int a = 0; a < D; a++) {
                real s = 0.0;
                for (int l = 0; l < D; l++)
                    s += (ddJdyy[i0][l] * Y[i1] + ddJdyT[l][i0] * Tv[i1]) * U[l][a];
                real v = -I * s;
                if (a == i1) v -= I * dJdy[i0];
                ddR[(1 + a) * P + 1 + D + k] = v;
            }
        }
        /* x-x block */
        for (int a = 0; a < D; a++)
            for (int b = a; b < D; b++) {
                real s = 0.0;
                for (int k = 0; k < D; k++)
                    for (int l = 0; l < D; l++) s += ddJdyy[k][l] * U[k][a] * U[l][b];
                ddR[(1 + a) * P + 1 + b] = I * s;
            }
        for (int i = 0; i < P; i++)
            for (int j = 0; j < i; j++) ddR[i * P + j] = ddR[j * P + i];
    }
    return 1;
}

/* Detector point -> volume point.
 * kind 3: y = p0 w0[a] + p1 w1[a]                (NewtonBases.py:676-680)
 * kind 25: y = (p0 w0[a,0], p0 w0[a,1], p1)      (NewtonBases.py:684-688)
 * kind 2: y = p0 w0[a]                            (NewtonBases2D.py:246-249) */
static inline void tovec(int kind, real p0, real p1, const real *w0, const real *w1,
                         real *y)
{
    if (kind == 3) {
        for (int i = 0; i < 3; i++) y[i] = p0 * w0[i] + p1 * w1[i];
    } else if (kind == 25) {
        y[0] = p0 * w0[0];
        y[1] = p0 * w0[1];
        y[2] = p1;
    } else {
        y[0] = p0 * w0[0];
        y[1] = p0 * w0[1];
    }
}

/* Geometry descriptor passed from Python.  lenT = 3 (3D), 2 (2.5D and 2D). */
static int geom_kind(int D, int lenT) { return D == 2 ? 2 : (lenT == 3 ? 3 : 25); }

/*
 * Forward projection R[a, i0, i1] = sum_atoms __Rf.
 * D: 2 or 3.  r: N x (D(D+1)/2) anisotropic slots.  T, w0: nA x lenT, w1: nA x 3 (3D only).
 * n1 = 1 for D = 2.
 */
void or_radproj(int D, int N, const real *I, const real *x, const real *r, int nA, int lenT,
                const real *T, const real *w0, const real *w1, int n0, const real *p0,
                int n1, const real *p1, real *R)
{
    const int nr = D * (D + 1) / 2, kind = geom_kind(D, lenT);
    real(*Us)[3][3] = malloc(sizeof(real[3][3]) * (N > 0 ? N : 1));
    for (int j = 0; j < N; j++) fill_U(D, r + (size_t)j * nr, Us[j]);
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < nA; a++)
        for (int k0 = 0; k0 < n0; k0++) {
            real T3[3] = {T[a * lenT], T[a * lenT + 1], lenT == 3 ? T[a * lenT + 2] : 0.0};
            for (int k1 = 0; k1 < n1; k1++) {
                real y[3];
                tovec(kind, p0[k0], D == 3 ? p1[k1] : 0.0, w0 + (size_t)a * lenT,
                      w1 ? w1 + (size_t)a * 3 : NULL, y);
                real s = 0.0;
                for (int j = 0; j < N; j++) s += atom_Rf(D, I[j], x + (size_t)j * D, Us[j], y, T3);
                R[((size_t)a * n0 + k0) * n1 + k1] = s;
            }
        }
    free(Us);
}

/*
 * Per-atom derivatives over all rays (batched over atoms, each atom independent).
 * mode 1 (L2):     f = 1/2 sum (R-C)^2, Df = sum (R-C) dR, DDf = sum dR dR^T + (R-C) ddR
 * mode 0 (linear): f = sum R C,         Df = sum C dR,     DDf = sum C ddR
 * NewtonBases.py:792-870 (L2), 993-1069 (linear); 2D at NewtonBases2D.py:329-401, 444-515.
 * Outputs f[N], Df[N*P], DDf[N*P*P].
 */
void or_derivs(int D, int mode, int order, int N, const real *I, const real *x, const real *r,
               int nA, int lenT, const real *T, const real *w0, const real *w1, int n0,
               const real *p0, int n1, const real *p1, const real *C, real *f, real *Df,
               real *DDf)
{
    const int nr = D * (D + 1) / 2, P = 1 + D + nr, kind = geom_kind(D, lenT);
#pragma omp parallel for schedule(dynamic, 1)
    for (int j = 0; j < N; j++) {
        real U[3][3];
        fill_U(D, r + (size_t)j * nr, U);
        real F = 0.0, dR[10], ddR[100];
        real *G = Df + (size_t)j * P, *H = DDf + (size_t)j * P * P;
        memset(G, 0, sizeof(real) * P);
        memset(H, 0, sizeof(real) * P * P);
        for (int a = 0; a < nA; a++) {
            real T3[3] = {T[a * lenT], T[a * lenT + 1], lenT == 3 ? T[a * lenT + 2] : 0.0};
            for (int k0 = 0; k0 < n0; k0++)
                for (int k1 = 0; k1 < n1; k1++) {
                    real y[3], Rv;
                    real c = C[((size_t)a * n0 + k0) * n1 + k1];
                    tovec(kind, p0[k0], D == 3 ? p1[k1] : 0.0, w0 + (size_t)a * lenT,
                          w1 ? w1 + (size_t)a * 3 : NULL, y);
                    int in = atom_dRf(D, I[j], x + (size_t)j * D, r + (size_t)j * nr, U, y, T3,
                                      order, &Rv, dR, ddR);
                    if (mode == 1) {
                        if (!in) {
                            F += c * c;
                            continue;
                        }
                        real res = Rv - c;
                        F += res * res;
                        for (int p = 0; p < P; p++) {
                            G[p] += res * dR[p];
                            for (int q = 0; q < P; q++)
                                H[p * P + q] += dR[p] * dR[q] + (order > 1 ? res * ddR[p * P + q] : 0.0);
                        }
                    } else {
                        if (!in) continue;
                        F += Rv * c;
                        for (int p = 0; p < P; p++) {
                            G[p] += c * dR[p];
                            if (order > 1)
                                for (int q = 0; q < P; q++) H[p * P + q] += c * ddR[p * P + q];
                        }
                    }
                }
        }
        f[j] = (mode == 1) ? 0.5 * F : F;
    }
}

/* Volume render u[j0,j1(,j2)] = sum_atoms __f.  NewtonBases.py:698-711, NewtonBases2D.py:259-271 */
void or_volproj(int D, int N, const real *I, const real *x, const real *r, int m0,
                const real *y0, int m1, const real *y1, int m2, const real *y2, real *u)
{
    const int nr = D * (D + 1) / 2;
    if (D == 2) m2 = 1;
    real(*Us)[3][3] = malloc(sizeof(real[3][3]) * (N > 0 ? N : 1));
    for (int j = 0; j < N; j++) fill_U(D, r + (size_t)j * nr, Us[j]);
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < m0; a++)
        for (int b = 0; b < m1; b++)
            for (int c = 0; c < m2; c++) {
                real y[3] = {y0[a], y1[b], D == 3 ? y2[c] : 0.0};
                real s = 0.0;
                for (int j = 0; j < N; j++) s += atom_f(D, I[j], x + (size_t)j * D, Us[j], y);
                u[((size_t)a * m1 + b) * m2 + c] = s;
            }
    free(Us);
}

/*
 * Fourier-space evaluation at explicit k-points K[npts x D] (grids and
 * hyperplanes are expanded to points by the Python side):
 *   3D aniso: sum I exp(-|U^T k|^2/2) exp(-i x.k) if |U^T k|^2 < 30   (FTGaussian3D.py:95-148)
 *   3D iso:   norm = |k|^2 / r^2, same cutoff                          (FTGaussian3D.py:151-196)
 *   2D aniso: U = [[r0, r2], [0, r1]], cutoff 60                       (FTGaussian2D.py:42-70)
 *   2D iso:   norm = r^2 |k|^2, cutoff 60                              (FTGaussian2D.py:73-100)
 * iso != 0 selects the isotropic form (r has one slot per atom).
 * out is interleaved complex (re, im).
 */
void or_ft_eval(int D, int iso, int N, const real *I, const real *x, const real *r,
                long npts, const real *K, real *out)
{
    const int nr = iso ? 1 : D * (D + 1) / 2;
    const real cut = (D == 3) ? 30.0 : 60.0;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < npts; p++) {
        const real *k = K + p * D;
        real re = 0.0, im = 0.0;
        for (int j = 0; j < N; j++) {
            const real *rr = r + (size_t)j * nr;
            real norm;
            if (iso) {
                real k2 = 0.0;
                for (int i = 0; i < D; i++) k2 += k[i] * k[i];
                norm = (D == 3) ? k2 / (rr[0] * rr[0]) : rr[0] * rr[0] * k2;
            } else {
                real U[3][3];
                fill_U(D, rr, U);
                norm = 0.0;
                for (int i = 0; i < D; i++) { /* (U^T k)_i = sum_{l<=i} U[l][i] k_l */
                    real s = 0.0;
                    for (int l = 0; l <= i; l++) s += U[l][i] * k[l];
                    norm += s * s;
                }
            }
            if (norm < cut) {
                real ip = 0.0;
                for (int i = 0; i < D; i++) ip += x[(size_t)j * D + i] * k[i];
                real amp = I[j] * exp(-0.5 * norm);
                re += amp * cos(ip);
                im -= amp * sin(ip);
            }
        }
        out[2 * p] = re;
        out[2 * p + 1] = im;
    }
}

/* Number of OpenMP threads the oracle runs with (reported by the CPU baseline). */
int or_num_threads(void)
{
#ifdef _OPENMP
    extern int omp_get_max_threads(void);
    return omp_get_max_threads();
#else
    return 1;
#endif
}
