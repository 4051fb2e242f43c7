// gt_reg.h -- per-atom regularisers of the fitters, FP64.
//
// GaussDictCode/regularisation.py:22-442 (null, Joubert, Mass, Volume,
// Radius, Iso).  Every regulariser there is a sum over atoms of one term per
// atom, so one atom's value, gradient and Hessian are all a fitter needs: a
// change of atom j changes the total by reg(new_j) - reg(old_j).
//
// Parameters are those of the solve: th = [I, x_0..x_{D-1}, r...] with one r
// for isotropic atoms and the D(D+1)/2 slots [r00, r11, (r22,) r01, (r12,
// r02)] otherwise.  The value is the reference's __call__ term for one atom,
// +inf outside its domain.  Gradient and Hessian are the exact derivatives
// of that value; where the reference's grad/hess are not (Volume.grad drops
// the factor k, Volume.hess is not the derivative and names an undefined
// attribute, Mass.hess leaves out some r_i r_j cross terms, Joubert.hess is
// empty) the derivative of the value is what is computed here.  Joubert's
// domain test covers the diagonal slots that enter its logarithm (the
// reference tests every slot, regularisation.py:63-64, which makes the value
// infinite whenever an off-diagonal slot is <= 0).
#pragma once

#include <hip/hip_runtime.h>

#include "gausstomo.h"

namespace gt {

// Number of solve parameters of one atom.
__host__ __device__ inline int reg_nparam(int D, int iso) { return 1 + D + (iso ? 1 : D * (D + 1) / 2); }

// The regulariser terms run inside the fit kernel, whose register budget is
// set by its pixel loops: the general FP64 pow and log stay out of line (a
// call on the regulariser's own path only) and integral exponents, the usual
// case (dimension powers, k = 2, 3), are repeated multiplications.
__device__ __noinline__ double reg_pow_general(double x, double y) { return pow(x, y); }
__device__ __noinline__ double reg_log(double x) { return log(x); }

__device__ inline double reg_ipow(double x, int n)
{
    double r = 1.0, b = n < 0 ? 1.0 / x : x;
    for (unsigned e = n < 0 ? -n : n; e; e >>= 1) {
        if (e & 1u) r *= b;
        b *= b;
    }
    return r;
}

__device__ inline double reg_pow(double x, double y)
{
    return (y == rint(y) && fabs(y) <= 64.0) ? reg_ipow(x, (int)y) : reg_pow_general(x, y);
}

// Value of the regulariser term of one atom; if g != nullptr also its
// gradient g[PI] and Hessian H[PI * PI] (row-major), both overwritten.
template <int D>
__device__ inline double reg_eval(int kind, const double *p, int iso, const double *th, double *g, double *H)
{
    constexpr int NR = D * (D + 1) / 2;
    const int nr = iso ? 1 : NR;
    const int PI = 1 + D + nr;
    const double INF = __builtin_huge_val();
    if (g) {
        for (int i = 0; i < PI; i++) g[i] = 0.0;
        for (int i = 0; i < PI * PI; i++) H[i] = 0.0;
    }
    const double I = th[0];
    const double *x = th + 1;
    const double *r = th + 1 + D;
    const int o = 1 + D;  // index of the first r parameter
    double diag[3] = {1.0, 1.0, 1.0};
    double pd = 1.0;  // product of the diagonal slots (aniso)
    bool diag_le0 = false;
    for (int i = 0; i < D; i++) {
        diag[i] = iso ? r[0] : r[i];
        if (!iso) pd *= r[i];
        if (diag[i] <= 0.0) diag_le0 = true;
    }
    switch (kind) {
    case GT_REG_JOUBERT: {
        // -a log I + b/2 |x|^2 + c v - d log v - e,  v = r^D (iso) or prod(diag)
        // (regularisation.py:53-70, params from :41-51)
        const double a = p[0], b = p[1], c = p[2], d = p[3], e = p[4];
        const double v = iso ? reg_pow(r[0], (double)D) : pd;
        double xx = 0.0;
        for (int i = 0; i < D; i++) xx += x[i] * x[i];
        const double val = -a * reg_log(I) + 0.5 * b * xx + c * v - d * reg_log(v) - e;
        if (g) {
            g[0] = -a / I;
            H[0] = a / (I * I);
            for (int i = 0; i < D; i++) {
                g[1 + i] = b * x[i];
                H[(1 + i) * PI + 1 + i] = b;
            }
            if (iso) {
                const double rr = r[0];
                g[o] = D * c * reg_pow(rr, (double)(D - 1)) - D * d / rr;
                H[o * PI + o] = c * D * (D - 1) * reg_pow(rr, (double)(D - 2)) + d * D / (rr * rr);
            } else {
                for (int i = 0; i < D; i++) {
                    g[o + i] = (c * pd - d) / r[i];
                    for (int l = 0; l < D; l++)
                        H[(o + i) * PI + o + l] = (i == l) ? d / (r[i] * r[i]) : c * pd / (r[i] * r[l]);
                }
            }
        }
        return (I <= 0.0 || diag_le0) ? INF : val;
    }
    case GT_REG_MASS: {
        // s (|I R| - m)^k,  R = r^D (iso) or 1/prod(diag) (regularisation.py:114-129)
        const double m = p[0], k = p[1], s = p[2];
        const double R = iso ? reg_pow(r[0], (double)D) : 1.0 / pd;
        const double V = I * R;
        const double val = s * reg_pow(fabs(V) - m, k);
        if (g) {
            // derivatives of (|V| - m)^k: first-order factor with the sign of V
            // (taken as + at V = 0, the reference's grad formula :131-160)
            const double sg = V < 0.0 ? -1.0 : 1.0, Va = fabs(V);
            const double s0 = sg * k * reg_pow(Va - m, k - 1.0), s1 = k * (k - 1.0) * reg_pow(Va - m, k - 2.0);
            g[0] = s * s0 * R;
            H[0] = s * R * R * s1;
            if (iso) {
                const double rr = r[0];
                const double dR = D * reg_pow(rr, (double)(D - 1));
                const double ddR = D * (D - 1) * reg_pow(rr, (double)(D - 2));
                g[o] = s * s0 * I * dR;
                H[o] = H[o * PI] = s * dR * (s0 + R * I * s1);
                H[o * PI + o] = s * (I * ddR * s0 + (I * dR) * (I * dR) * s1);
            } else {
                for (int i = 0; i < D; i++) {
                    g[o + i] = -s * s0 * I * R / r[i];
                    H[o + i] = H[(o + i) * PI] = -s * R / r[i] * (s0 + I * R * s1);
                    for (int l = 0; l < D; l++)
                        H[(o + i) * PI + o + l] = s * I * R / (r[i] * r[l]) * ((i == l ? 2.0 : 1.0) * s0 + I * R * s1);
                }
            }
        }
        return (I < 0.0) ? INF : val;
    }
    case GT_REG_VOLUME: {
        // s |d|^k,  d = r^-D - r0^D (iso) or prod(diag) - r0^D (regularisation.py:218-236)
        const double p0 = p[0], k = p[1], s = p[2];
        const double dd = iso ? reg_pow(r[0], -(double)D) - p0 : pd - p0;
        const double val = s * reg_pow(fabs(dd), k);
        if (g) {
            const double f1 = s * k * dd * reg_pow(fabs(dd), k - 2.0);
            const double f2 = s * k * (k - 1.0) * reg_pow(fabs(dd), k - 2.0);
            if (iso) {
                const double rr = r[0];
                const double d1 = -D * reg_pow(rr, -(double)D - 1.0);
                const double d2 = D * (D + 1) * reg_pow(rr, -(double)D - 2.0);
                g[o] = f1 * d1;
                H[o * PI + o] = f2 * d1 * d1 + f1 * d2;
            } else {
                for (int i = 0; i < D; i++) {
                    g[o + i] = f1 * pd / r[i];
                    for (int l = 0; l < D; l++)
                        H[(o + i) * PI + o + l] =
                            f2 * (pd / r[i]) * (pd / r[l]) + ((i == l) ? 0.0 : f1 * pd / (r[i] * r[l]));
                }
            }
        }
        return diag_le0 ? INF : val;
    }
    case GT_REG_RADIUS: {
        // s sum_i (r_i - r0)^kk over the diagonal, iso D (r - r0)^kk; kk = -k
        // (regularisation.py:339-350)
        const double r0 = p[0], s = p[1], kk = p[2];
        bool below = false;
        double val = 0.0;
        if (iso) {
            const double t = r[0] - r0;
            below = r[0] <= r0;
            val = s * D * reg_pow(t, kk);
            if (g) {
                g[o] = s * D * kk * reg_pow(t, kk - 1.0);
                H[o * PI + o] = s * D * kk * (kk - 1.0) * reg_pow(t, kk - 2.0);
            }
        } else {
            for (int i = 0; i < D; i++) {
                const double t = r[i] - r0;
                if (r[i] <= r0) below = true;
                val += reg_pow(t, kk);
                if (g) {
                    g[o + i] = s * kk * reg_pow(t, kk - 1.0);
                    H[(o + i) * PI + o + i] = s * kk * (kk - 1.0) * reg_pow(t, kk - 2.0);
                }
            }
            val *= s;
        }
        return below ? INF : val;
    }
    case GT_REG_ISO: {
        // s/2 sum of the squared off-diagonal slots (regularisation.py:398-407)
        const double s = p[0];
        if (iso) return 0.0;
        double val = 0.0;
        for (int k = D; k < NR; k++) {
            val += r[k] * r[k];
            if (g) {
                g[o + k] = s * r[k];
                H[(o + k) * PI + o + k] = s;
            }
        }
        return 0.5 * s * val;
    }
    default:
        return 0.0;
    }
}

}  // namespace gt
