// gt_pair.h -- per-(atom, angle) FP64 algebra of the Gaussian-atom projector.
//
// The reference evaluates, for every ray, the line integral of one atom
// (GaussDictCode/bin/NewtonBases.py:260-310) and its first and second
// derivatives with respect to theta = [I, x, r] (NewtonBases.py:313-527).
// For a fixed atom and a fixed projection angle that line integral is a 2D
// (1D for 2D problems) Gaussian on the detector:
//
//   R(p) = alpha * exp(-30 m(p)) * [m(p) <= 1],   alpha = I / |U T|,
//   m(p) = d^T B d,   d = (p - p*, 1),
//
// where U is the upper-triangular atom matrix (r slots, NewtonBases.py:249-251,
// 319-325), T the ray direction, y = G p the detector-to-volume map
// (NewtonBases.py:676-688, NewtonBases2D.py:246-249) and
//
//   B_ab = h_a^T Q h_b,  Q = U^T (Id - t t^T / s) U,  t = U T,  s = |t|^2,
//   h_k = G e_k (k < K),  h_K = G p* - x.
//
// This identity is exact (the reference's m = |(Id - t^ t^T) U (y - x)|^2
// expands to it), so the projector is: one FP64 setup per (atom, angle) here,
// then one quadratic, one compare and one exp2 per pixel in FP32.
//
// Derivatives use the same structure.  R depends on theta only through the
// intermediate vector gamma = (alpha, B_ab [a <= b]), and R is a polynomial in
// d times exp(-30 m), so every pixel sum the reference accumulates
// (NewtonBases.py:833-839, 1034-1039) is a contraction of the pixel moments
//
//   S1[mono] = sum_p w1(p) mono(p - p*),   S2[mono] = sum_p E(p)^2 mono(p - p*)
//
// over the monomials of degree <= 4 (15 in 2D detectors, 5 in 1D), with
// w1 = (lambda alpha E - sigma C) E chosen per mode.  pair_finish() turns those
// moments into this angle's contribution to (f, grad, Hessian) with the exact
// chain rule gamma -> theta, in FP64.
#pragma once

#include <hip/hip_runtime.h>

namespace gt {

// -30 * log2(e): m -> exponent of exp2
constexpr double KAPPA = -43.28085122666891;
constexpr int MAXD = 3;

// r-slot -> (row, col) of U.  NewtonBases.py:319-325, NewtonBases2D.py:84-87.
__host__ __device__ inline int slot_row(int D, int k)
{
    return D == 3 ? (k < 3 ? k : (k == 3 ? 0 : (k == 4 ? 1 : 0))) : (k < 2 ? k : 0);
}
__host__ __device__ inline int slot_col(int D, int k)
{
    return D == 3 ? (k < 3 ? k : (k == 3 ? 1 : 2)) : (k < 2 ? k : 1);
}

template <int D> struct Dims {
    static constexpr int NR = D * (D + 1) / 2;  // r slots
    static constexpr int P = 1 + D + NR;        // parameters [I, x, r]
    static constexpr int K = D - 1;             // detector dims
    static constexpr int NB = (K + 1) * (K + 2) / 2;  // B entries a <= b over {0..K}
    static constexpr int NG = 1 + NB;           // gamma = (alpha, B)
    static constexpr int NM = (K == 2) ? 15 : 5;  // monomials of degree <= 4
};

// Monomial index of dp0^i dp1^j (K = 2) or dp^i (K = 1).
__host__ __device__ inline int mono2(int i, int j)
{
    int deg = i + j;
    return deg * (deg + 1) / 2 + j;
}

// B-pair index p -> (a, b), a <= b, over {0..K}.
template <int K> __host__ __device__ inline void bpair(int p, int &a, int &b)
{
    if (K == 2) {
        const int A[6] = {0, 0, 0, 1, 1, 2}, Bv[6] = {0, 1, 2, 1, 2, 2};
        a = A[p];
        b = Bv[p];
    } else {
        const int A[3] = {0, 0, 1}, Bv[3] = {0, 1, 1};
        a = A[p];
        b = Bv[p];
    }
}

// Exponents (of dp0, dp1) of d_a: d = (dp0, dp1, 1) for K = 2, (dp, 1) for K = 1.
template <int K> __host__ __device__ inline void dexp(int a, int &e0, int &e1)
{
    e0 = (a == 0 && K >= 1) ? 1 : 0;
    e1 = (K == 2 && a == 1) ? 1 : 0;
}

template <int K> __host__ __device__ inline int mono_index(int e0, int e1)
{
    return K == 2 ? mono2(e0, e1) : e0;
}

// Geometry of one (atom, angle) pair in FP64.
template <int D> struct PairGeom {
    double U[3][3];
    double T[3];
    double G[2][3];  // detector basis columns g_k: y = sum_k p_k g_k
    double x[3];
    double I;
};

// Quantities shared by setup and finish.
template <int D> struct PairAlg {
    double t[3], s, alpha;
    double h[3][3];   // h_0..h_K
    double u[3][3];   // U h_a
    double mu[3];     // (u_a . t) / s
    double ut[3][3];  // M u_a
    double ht[3][3];  // h_a - mu_a T
    double B[3][3];   // B_ab
};

template <int D> __host__ __device__ inline void matvec_upper(const double U[3][3], const double *v, double *out)
{
    for (int i = 0; i < D; i++) {
        double acc = 0.0;
        for (int j = i; j < D; j++) acc += U[i][j] * v[j];
        out[i] = acc;
    }
}

// Fill PairAlg at expansion point pstar.
template <int D>
__host__ __device__ inline void pair_alg(const PairGeom<D> &g, const double *pstar, PairAlg<D> &a)
{
    constexpr int K = Dims<D>::K;
    matvec_upper<D>(g.U, g.T, a.t);
    a.s = 0.0;
    for (int i = 0; i < D; i++) a.s += a.t[i] * a.t[i];
    a.alpha = g.I / sqrt(a.s);
    for (int k = 0; k < K; k++)
        for (int i = 0; i < D; i++) a.h[k][i] = g.G[k][i];
    for (int i = 0; i < D; i++) {
        double y = 0.0;
        for (int k = 0; k < K; k++) y += pstar[k] * g.G[k][i];
        a.h[K][i] = y - g.x[i];
    }
    for (int b = 0; b <= K; b++) {
        matvec_upper<D>(g.U, a.h[b], a.u[b]);
        double ip = 0.0;
        for (int i = 0; i < D; i++) ip += a.u[b][i] * a.t[i];
        a.mu[b] = ip / a.s;
        for (int i = 0; i < D; i++) {
            a.ut[b][i] = a.u[b][i] - a.mu[b] * a.t[i];
            a.ht[b][i] = a.h[b][i] - a.mu[b] * g.T[i];
        }
    }
    for (int b0 = 0; b0 <= K; b0++)
        for (int b1 = 0; b1 <= K; b1++) {
            double v = 0.0;
            for (int i = 0; i < D; i++) v += a.u[b0][i] * a.ut[b1][i];
            a.B[b0][b1] = v;
        }
}

// FP32 record of one (atom, angle) pair, consumed by the pixel kernels.
struct alignas(16) PairRec {
    float alpha;
    float c_hi[2], c_lo[2];  // expansion point p* = c_hi + c_lo
    float q[6];              // kappa * m coefficients (see pair_q)
    float pad;
    int lo[2], hi[2];        // inclusive detector index box, lo > hi if empty
};
static_assert(sizeof(PairRec) == 64, "PairRec must stay 64 bytes");

// Index searches on an ascending coordinate vector p of length n.  The
// first guess is read off the end points (exact on a uniform grid) and kept
// when its neighbours confirm it, so a uniform detector costs two dependent
// loads instead of a bisection's log2(n); a guess that fails the check falls
// back to bisection, so the result is the same for every ascending vector.

// First index i with p[i] >= v, n if none (0 for v NaN).
template <typename T> __host__ __device__ inline int lower_idx(const T *p, int n, double v)
{
    if (n <= 0) return 0;
    const double a = p[0], b = p[n - 1];
    if (!(v > a)) return 0;
    if (v > b) return n;
    // here a < v <= b, so n >= 2 and b > a
    int g = (int)ceil((v - a) / (b - a) * (double)(n - 1));
    g = g < 1 ? 1 : (g > n - 1 ? n - 1 : g);
    if ((double)p[g] >= v && (double)p[g - 1] < v) return g;
    int lo = 1, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((double)p[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}
// Last index i with p[i] <= v, -1 if none (-1 for v NaN).
template <typename T> __host__ __device__ inline int upper_idx(const T *p, int n, double v)
{
    if (n <= 0) return -1;
    const double a = p[0], b = p[n - 1];
    if (!(v >= a)) return -1;
    if (v >= b) return n - 1;
    // here a <= v < b, so n >= 2 and b > a
    int g = (int)floor((v - a) / (b - a) * (double)(n - 1));
    g = g < 0 ? 0 : (g > n - 2 ? n - 2 : g);
    if ((double)p[g] <= v && (double)p[g + 1] > v) return g;
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if ((double)p[mid] <= v) lo = mid + 1;
        else hi = mid;
    }
    return lo - 1;
}

// Setup of one pair: expansion point, FP32 quadratic and the detector box.
// det[k], nd[k]: detector coordinate vectors; cull == 0 disables the box.
template <int D>
__host__ __device__ inline void pair_setup(const PairGeom<D> &g, const float *const *det, const int *nd,
                                           int cull, PairRec &rec)
{
    constexpr int K = Dims<D>::K;
    // Quadratic in p around the origin: m(p) = p^T A p + 2 b^T p + c.
    double zero[2] = {0.0, 0.0};
    PairAlg<D> a0;
    pair_alg<D>(g, zero, a0);
    double pstar[2] = {0.0, 0.0};
    bool degenerate = false;
    if (K == 2) {
        double A00 = a0.B[0][0], A01 = a0.B[0][1], A11 = a0.B[1][1];
        double b0 = a0.B[0][2], b1 = a0.B[1][2];
        double det2 = A00 * A11 - A01 * A01;
        if (det2 > 1e-12 * A00 * A11 && A00 > 0.0 && A11 > 0.0) {
            pstar[0] = -(A11 * b0 - A01 * b1) / det2;
            pstar[1] = -(A00 * b1 - A01 * b0) / det2;
        } else {
            degenerate = true;
        }
    } else {
        double A = a0.B[0][0], b = a0.B[0][1];
        if (A > 0.0) pstar[0] = -b / A;
        else degenerate = true;
    }
    PairAlg<D> a;
    pair_alg<D>(g, pstar, a);
    rec.alpha = (float)a.alpha;
    for (int k = 0; k < 2; k++) {
        float hi = (float)pstar[k];
        rec.c_hi[k] = hi;
        rec.c_lo[k] = (float)(pstar[k] - (double)hi);
    }
    const double kp = KAPPA;
    for (int i = 0; i < 6; i++) rec.q[i] = 0.0f;
    if (K == 2) {
        rec.q[0] = (float)(kp * a.B[0][0]);
        rec.q[1] = (float)(2.0 * kp * a.B[0][1]);
        rec.q[2] = (float)(kp * a.B[1][1]);
        rec.q[3] = (float)(2.0 * kp * a.B[0][2]);
        rec.q[4] = (float)(2.0 * kp * a.B[1][2]);
        rec.q[5] = (float)(kp * a.B[2][2]);
    } else {
        rec.q[0] = (float)(kp * a.B[0][0]);
        rec.q[3] = (float)(2.0 * kp * a.B[0][1]);
        rec.q[5] = (float)(kp * a.B[1][1]);
    }
    rec.pad = 0.0f;
    // Detector box of {m <= 1}.
    for (int k = 0; k < 2; k++) {
        rec.lo[k] = 0;
        rec.hi[k] = (k < K) ? nd[k] - 1 : 0;
    }
    if (!cull || degenerate) return;
    double ctr[2], hw[2];
    bool empty = false;
    if (K == 2) {
        double A00 = a.B[0][0], A01 = a.B[0][1], A11 = a.B[1][1];
        double det2 = A00 * A11 - A01 * A01;
        double i00 = A11 / det2, i01 = -A01 / det2, i11 = A00 / det2;
        double b0 = a.B[0][2], b1 = a.B[1][2];
        double c0 = -(i00 * b0 + i01 * b1), c1 = -(i01 * b0 + i11 * b1);
        double rho2 = 1.0 - a.B[2][2] + (b0 * (i00 * b0 + i01 * b1) + b1 * (i01 * b0 + i11 * b1));
        if (!(rho2 >= 0.0)) empty = true;
        ctr[0] = pstar[0] + c0;
        ctr[1] = pstar[1] + c1;
        hw[0] = sqrt(fmax(rho2, 0.0) * i00);
        hw[1] = sqrt(fmax(rho2, 0.0) * i11);
    } else {
        double A = a.B[0][0], b = a.B[0][1];
        double rho2 = 1.0 - a.B[1][1] + b * b / A;
        if (!(rho2 >= 0.0)) empty = true;
        ctr[0] = pstar[0] - b / A;
        hw[0] = sqrt(fmax(rho2, 0.0) / A);
    }
    for (int k = 0; k < K; k++) {
        double pad = 1e-5 * hw[k] + 1e-7;
        if (empty || !(hw[k] == hw[k])) {
            rec.lo[k] = 1;
            rec.hi[k] = 0;
            continue;
        }
        rec.lo[k] = lower_idx(det[k], nd[k], ctr[k] - hw[k] - pad);
        rec.hi[k] = upper_idx(det[k], nd[k], ctr[k] + hw[k] + pad);
    }
}

// FP32 exponent q = kappa * m at offset (dp0, dp1).
__host__ __device__ inline float pair_q(const float *q, float dp0, float dp1)
{
    float row = fmaf(q[0], dp0, fmaf(q[1], dp1, q[3]));
    float col = fmaf(fmaf(q[2], dp1, q[4]), dp1, q[5]);
    return fmaf(row, dp0, col);
}

// Moment-to-derivative contraction for one pair, in FP64.
//   S1[NM]: sum w1 * mono,  S2[NM]: sum E^2 * mono (used when gn != 0)
//   The pixel weight is w1 = (lambda alpha E - sigma C) E; the finishing
//   only needs to know that v = sum res dR/dgamma uses w1 with "res" the
//   residual (L2, residual modes) or C itself (linear mode).
// Accumulates into grad[P] and, through the sink hs, the Hessian: hs.add(i,
// j, v) with i <= j adds v to entry (i, j) of the symmetric P x P matrix.
// Qgn (optional, replaces S2): the products sum_k M_k[m] M_k[m'] over the
// low monomials m, m' of degree <= 2 (NL = 6 for 2D detectors, 3 for 1D),
// for values that are pixel integrals M_k[m] = int_k mono E rather than point
// samples (the Gauss-Newton term is then not a single moment).
template <int P> struct HessFull {  // full P x P array, both triangles
    double *h;
    __host__ __device__ void add(int i, int j, double v) const
    {
        h[i * P + j] += v;
        if (i != j) h[j * P + i] += v;
    }
};
template <int P> struct HessUpper {  // upper triangle by rows: (0,0..P-1), (1,1..P-1), ...
    double *h;
    __host__ __device__ void add(int i, int j, double v) const { h[i * P - i * (i - 1) / 2 + (j - i)] += v; }
};

template <int D, class Sink>
__host__ __device__ inline void pair_finish_to(const PairGeom<D> &g, const double *pstar, const double *S1,
                                               const double *S2, int gn, int want_hess, double *grad, const Sink &hs,
                                               const double *Qgn = nullptr)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, NB = Dm::NB, NG = Dm::NG, P = Dm::P, NR = Dm::NR;
    PairAlg<D> a;
    pair_alg<D>(g, pstar, a);
    const double al = a.alpha, s = a.s;

    // psi_p = (2 - delta_ab) d_a d_b: factor and exponents
    double fac[NB];
    int pe0[NB], pe1[NB];
#pragma unroll
    for (int p = 0; p < NB; p++) {
        int x0, x1, ea0, ea1, eb0, eb1;
        bpair<K>(p, x0, x1);
        dexp<K>(x0, ea0, ea1);
        dexp<K>(x1, eb0, eb1);
        fac[p] = (x0 == x1) ? 1.0 : 2.0;
        pe0[p] = ea0 + eb0;
        pe1[p] = ea1 + eb1;
    }
    // v = sum res dR/dgamma ; Kg = sum dR dR^T (gn) + sum res d2R
    double v[NG], Kg[NG][NG];
    constexpr int NL = (K == 2) ? 6 : 3;
    v[0] = S1[0];
#pragma unroll
    for (int p = 0; p < NB; p++) v[1 + p] = -30.0 * al * fac[p] * S1[mono_index<K>(pe0[p], pe1[p])];
    Kg[0][0] = gn ? (Qgn ? Qgn[0] : S2[0]) : 0.0;
#pragma unroll
    for (int p = 0; p < NB; p++) {
        int m = mono_index<K>(pe0[p], pe1[p]);
        const double gnv = gn ? (Qgn ? Qgn[m] : S2[m]) : 0.0;
        double val = -30.0 * fac[p] * (S1[m] + al * gnv);
        Kg[0][1 + p] = val;
        Kg[1 + p][0] = val;
    }
#pragma unroll
    for (int p = 0; p < NB; p++)
#pragma unroll
        for (int q = p; q < NB; q++) {
            int m = mono_index<K>(pe0[p] + pe0[q], pe1[p] + pe1[q]);
            const double gnv =
                gn ? (Qgn ? Qgn[mono_index<K>(pe0[p], pe1[p]) * NL + mono_index<K>(pe0[q], pe1[q])] : S2[m]) : 0.0;
            double val = 900.0 * fac[p] * fac[q] * (al * S1[m] + al * al * gnv);
            Kg[1 + p][1 + q] = val;
            Kg[1 + q][1 + p] = val;
        }

    // Jacobian J[NG][P] of gamma w.r.t. theta = [I, x, r].
    double J[NG][P];
    for (int i = 0; i < NG; i++)
        for (int j = 0; j < P; j++) J[i][j] = 0.0;
    double rs = 1.0 / sqrt(s), rs3 = rs / s;
    J[0][0] = rs;
    for (int k = 0; k < NR; k++) {
        int i = slot_row(D, k), j = slot_col(D, k);
        J[0][1 + D + k] = -g.I * rs3 * a.t[i] * g.T[j];
    }
    // Q h_a = U^T (M u_a)
    double Qh[3][3];
    for (int b = 0; b <= K; b++)
        for (int c = 0; c < D; c++) {
            double acc = 0.0;
            for (int i = 0; i <= c; i++) acc += g.U[i][c] * a.ut[b][i];
            Qh[b][c] = acc;
        }
    for (int p = 0; p < NB; p++) {
        int x0, x1;
        bpair<K>(p, x0, x1);
        for (int c = 0; c < D; c++) {
            double val = 0.0;
            if (x0 == K) val -= Qh[x1][c];
            if (x1 == K) val -= Qh[x0][c];
            J[1 + p][1 + c] = val;
        }
        for (int k = 0; k < NR; k++) {
            int i = slot_row(D, k), j = slot_col(D, k);
            J[1 + p][1 + D + k] = a.ut[x0][i] * a.ht[x1][j] + a.ut[x1][i] * a.ht[x0][j];
        }
    }

    // grad += J^T v
    for (int j = 0; j < P; j++) {
        double acc = 0.0;
        for (int i = 0; i < NG; i++) acc += J[i][j] * v[i];
        grad[j] += acc;
    }
    if (!want_hess) return;
    // hess += J^T Kg J
    double KJ[NG][P];
    for (int i = 0; i < NG; i++)
        for (int j = 0; j < P; j++) {
            double acc = 0.0;
            for (int l = 0; l < NG; l++) acc += Kg[i][l] * J[l][j];
            KJ[i][j] = acc;
        }
    for (int i = 0; i < P; i++)
        for (int j = i; j < P; j++) {
            double acc = 0.0;
            for (int l = 0; l < NG; l++) acc += J[l][i] * KJ[l][j];
            hs.add(i, j, acc);
        }

    // hess += Hessian of Phi(theta) = v_alpha alpha + sum_ab V_ab B_ab  (v fixed)
    double V[3][3];
    for (int p = 0; p < NB; p++) {
        int x0, x1;
        bpair<K>(p, x0, x1);
        double w = (x0 == x1) ? v[1 + p] : 0.5 * v[1 + p];
        V[x0][x1] = w;
        V[x1][x0] = w;
    }
    double Hh[3][3], Uh[3][3], Uu[3][3];
    for (int i = 0; i < D; i++)
        for (int j = 0; j < D; j++) {
            double hh = 0.0, uh = 0.0, uu = 0.0;
            for (int b0 = 0; b0 <= K; b0++)
                for (int b1 = 0; b1 <= K; b1++) {
                    hh += V[b0][b1] * a.ht[b0][i] * a.ht[b1][j];
                    uh += V[b0][b1] * a.ut[b0][i] * a.ht[b1][j];
                    uu += V[b0][b1] * a.ut[b0][i] * a.ut[b1][j];
                }
            Hh[i][j] = hh;
            Uh[i][j] = uh;
            Uu[i][j] = uu;
        }
    const double va = v[0];
    const double rs5 = rs3 / s;
    // I - r
    for (int k = 0; k < NR; k++) {
        int i = slot_row(D, k), j = slot_col(D, k);
        double val = -va * rs3 * a.t[i] * g.T[j];
        hs.add(0, 1 + D + k, val);
    }
    // x - x : 2 V_KK Q_cd, Q_cd = (U e_c) . M (U e_d)
    double z[3];  // U^T t
    for (int c = 0; c < D; c++) {
        double acc = 0.0;
        for (int i = 0; i <= c; i++) acc += g.U[i][c] * a.t[i];
        z[c] = acc;
    }
    double MUc[3][3];  // column c of M U
    for (int c = 0; c < D; c++)
        for (int i = 0; i < D; i++) MUc[c][i] = (i <= c ? g.U[i][c] : 0.0) - z[c] / s * a.t[i];
    for (int c = 0; c < D; c++)
        for (int d = c; d < D; d++) {  // Q is symmetric: one triangle
            double Qcd = 0.0;
            for (int i = 0; i <= c; i++) Qcd += g.U[i][c] * MUc[d][i];
            hs.add(1 + c, 1 + d, 2.0 * V[K][K] * Qcd);
        }
    // x - r
    double qt[3], uq[3];
    for (int i = 0; i < D; i++) {
        double a1 = 0.0, a2 = 0.0;
        for (int b = 0; b <= K; b++) {
            a1 += V[b][K] * a.ht[b][i];
            a2 += V[b][K] * a.ut[b][i];
        }
        qt[i] = a1;
        uq[i] = a2;
    }
    for (int c = 0; c < D; c++)
        for (int k = 0; k < NR; k++) {
            int i = slot_row(D, k), j = slot_col(D, k);
            double ec_j = (j == c ? 1.0 : 0.0) - z[c] / s * g.T[j];
            double val = -2.0 * (MUc[c][i] * qt[j] + uq[i] * ec_j);
            hs.add(1 + c, 1 + D + k, val);
        }
    // r - r
    for (int k = 0; k < NR; k++) {
        int i1 = slot_row(D, k), j1 = slot_col(D, k);
        for (int l = k; l < NR; l++) {
            int i2 = slot_row(D, l), j2 = slot_col(D, l);
            double dik = (i1 == i2) ? 1.0 : 0.0;
            double val = 2.0 * (dik * Hh[j2][j1] -
                                (a.t[i1] * a.t[i2] * Hh[j2][j1] + a.t[i1] * g.T[j2] * Uh[i2][j1] +
                                 g.T[j1] * a.t[i2] * Uh[i1][j2] + g.T[j1] * g.T[j2] * Uu[i1][i2]) / s);
            val += va * g.I *
                   (3.0 * rs5 * a.t[i1] * g.T[j1] * a.t[i2] * g.T[j2] - rs3 * dik * g.T[j1] * g.T[j2]);
            hs.add(1 + D + k, 1 + D + l, val);
        }
    }
}

// pair_finish_to with the Hessian in a full P x P array (both triangles).
template <int D>
__host__ __device__ inline void pair_finish(const PairGeom<D> &g, const double *pstar, const double *S1,
                                            const double *S2, int gn, int want_hess, double *grad, double *hess,
                                            const double *Qgn = nullptr)
{
    pair_finish_to<D>(g, pstar, S1, S2, gn, want_hess, grad, HessFull<Dims<D>::P>{hess}, Qgn);
}

// The same contraction split in two, for callers that spread one pair's
// (gradient, Hessian) entries over many threads: pair_prepare() stores every
// per-pair ingredient (Jacobian J of gamma, the gamma-space matrix Kg, the
// weights v and the pieces of the Hessian of Phi = v . gamma(theta)) in a
// PairFin, typically in LDS; fin_grad()/fin_hess() then produce one entry each.
// The algebra is identical to pair_finish().
template <int D> struct PairFin {
    static constexpr int P = Dims<D>::P, NG = Dims<D>::NG;
    double J[NG][P];
    double Kg[NG][NG];
    double KJ[NG][P];  // Kg J
    double v[NG];
    double va, I, s, rs3, rs5, VKK;
    double t[3], T[3], z[3], qt[3], uq[3];
    double MUc[3][3], Hh[3][3], Uh[3][3], Uu[3][3], Q[3][3];
    double pad;  // odd number of doubles: lanes reading consecutive PairFins spread over LDS banks
};
static_assert((sizeof(PairFin<3>) / 8) % 2 == 1 && (sizeof(PairFin<2>) / 8) % 2 == 1,
              "PairFin must hold an odd number of doubles");

// Per-entry pieces of pair_prepare, so that the ingredients of one pair can be
// formed by many lanes at once.  prep1 (v, Kg, J, z) needs PairAlg and the
// moments; prep2 (MUc, scalars) needs prep1; prep3 (Hh, Uh, Uu, Q, qt, uq)
// needs prep2.  Entry idx of each stage is independent of the others.
template <int D> struct PrepCount {
    static constexpr int NG = Dims<D>::NG, P = Dims<D>::P;
    static constexpr int N1 = NG + NG * NG + NG * P + 3;
    static constexpr int N2 = 10;
    static constexpr int N3 = 42;
    static constexpr int N4 = NG * P;
};

// B-pair index of (b0, b1) in any order.
template <int K> __host__ __device__ inline int bpair_index(int b0, int b1)
{
    const int a = b0 < b1 ? b0 : b1, b = b0 < b1 ? b1 : b0;
    return a * (K + 1) - a * (a - 1) / 2 + (b - a);
}

// V[b0][b1] of Phi's gamma weights (symmetric; off-diagonal entries halved).
template <int D> __host__ __device__ inline double vmat(const PairFin<D> &F, int b0, int b1)
{
    const double v = F.v[1 + bpair_index<Dims<D>::K>(b0, b1)];
    return (b0 == b1) ? v : 0.5 * v;
}

template <int D>
__host__ __device__ inline void prep1(const PairGeom<D> &g, const PairAlg<D> &a, const double *S1, const double *S2,
                                      int gn, int idx, PairFin<D> &F)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, NG = Dm::NG, P = Dm::P;
    const double al = a.alpha;
    if (idx < NG) {  // v
        if (idx == 0) {
            F.v[0] = S1[0];
            return;
        }
        int x0, x1, ea0, ea1, eb0, eb1;
        bpair<K>(idx - 1, x0, x1);
        dexp<K>(x0, ea0, ea1);
        dexp<K>(x1, eb0, eb1);
        const double fp = (x0 == x1) ? 1.0 : 2.0;
        F.v[idx] = -30.0 * al * fp * S1[mono_index<K>(ea0 + eb0, ea1 + eb1)];
        return;
    }
    idx -= NG;
    if (idx < NG * NG) {  // Kg
        const int l = idx / NG, m = idx - l * NG;
        if (l == 0 && m == 0) {
            F.Kg[0][0] = gn ? S2[0] : 0.0;
            return;
        }
        int x0, x1, ea0, ea1, eb0, eb1;
        bpair<K>((l == 0 ? m : l) - 1, x0, x1);
        dexp<K>(x0, ea0, ea1);
        dexp<K>(x1, eb0, eb1);
        const double fp = (x0 == x1) ? 1.0 : 2.0;
        if (l == 0 || m == 0) {
            const int mi = mono_index<K>(ea0 + eb0, ea1 + eb1);
            F.Kg[l][m] = -30.0 * fp * (S1[mi] + (gn ? al * S2[mi] : 0.0));
            return;
        }
        int y0, y1, fa0, fa1, fb0, fb1;
        bpair<K>(m - 1, y0, y1);
        dexp<K>(y0, fa0, fa1);
        dexp<K>(y1, fb0, fb1);
        const double fq = (y0 == y1) ? 1.0 : 2.0;
        const int mm = mono_index<K>(ea0 + eb0 + fa0 + fb0, ea1 + eb1 + fa1 + fb1);
        F.Kg[l][m] = 900.0 * fp * fq * (al * S1[mm] + (gn ? al * al * S2[mm] : 0.0));
        return;
    }
    idx -= NG * NG;
    if (idx < NG * P) {  // J = d gamma / d theta
        const int row = idx / P, j = idx - row * P;
        double val = 0.0;
        if (row == 0) {
            const double rs = 1.0 / sqrt(a.s), rs3 = rs / a.s;
            if (j == 0) {
                val = rs;
            } else if (j >= 1 + D) {
                const int k = j - 1 - D;
                val = -g.I * rs3 * a.t[slot_row(D, k)] * g.T[slot_col(D, k)];
            }
        } else {
            int x0, x1;
            bpair<K>(row - 1, x0, x1);
            if (j >= 1 && j <= D) {
                const int c = j - 1;
                if (x0 == K) {
                    double acc = 0.0;
                    for (int i = 0; i <= c; i++) acc += g.U[i][c] * a.ut[x1][i];
                    val -= acc;
                }
                if (x1 == K) {
                    double acc = 0.0;
                    for (int i = 0; i <= c; i++) acc += g.U[i][c] * a.ut[x0][i];
                    val -= acc;
                }
            } else if (j > D) {
                const int i = slot_row(D, j - 1 - D), jj = slot_col(D, j - 1 - D);
                val = a.ut[x0][i] * a.ht[x1][jj] + a.ut[x1][i] * a.ht[x0][jj];
            }
        }
        F.J[row][j] = val;
        return;
    }
    idx -= NG * P;
    if (idx < 3) {  // z = U^T t (upper U), t, T
        const int c = idx;
        double acc = 0.0;
        for (int i = 0; i <= c && c < D; i++) acc += g.U[i][c] * a.t[i];
        F.z[c] = acc;
        F.t[c] = a.t[c];
        F.T[c] = g.T[c];
    }
}

template <int D>
__host__ __device__ inline void prep2(const PairGeom<D> &g, const PairAlg<D> &a, int idx, PairFin<D> &F)
{
    constexpr int K = Dims<D>::K;
    if (idx < 9) {
        const int c = idx / 3, i = idx - 3 * c;
        F.MUc[c][i] = (c < D && i < D) ? ((i <= c ? g.U[i][c] : 0.0) - F.z[c] / a.s * a.t[i]) : 0.0;
    } else if (idx == 9) {
        const double s = a.s, rs = 1.0 / sqrt(s), rs3 = rs / s;
        F.va = F.v[0];
        F.I = g.I;
        F.s = s;
        F.rs3 = rs3;
        F.rs5 = rs3 / s;
        F.VKK = vmat<D>(F, K, K);
        F.pad = 0.0;
    }
}

template <int D>
__host__ __device__ inline void prep3(const PairGeom<D> &g, const PairAlg<D> &a, int idx, PairFin<D> &F)
{
    constexpr int K = Dims<D>::K;
    if (idx < 27) {
        const int which = idx / 9, i = (idx / 3) % 3, j = idx % 3;
        double acc = 0.0;
        if (i < D && j < D)
            for (int b0 = 0; b0 <= K; b0++)
                for (int b1 = 0; b1 <= K; b1++) {
                    const double V = vmat<D>(F, b0, b1);
                    if (which == 0)
                        acc += V * a.ht[b0][i] * a.ht[b1][j];
                    else if (which == 1)
                        acc += V * a.ut[b0][i] * a.ht[b1][j];
                    else
                        acc += V * a.ut[b0][i] * a.ut[b1][j];
                }
        (which == 0 ? F.Hh : which == 1 ? F.Uh : F.Uu)[i][j] = acc;
    } else if (idx < 36) {
        const int c = (idx - 27) / 3, d = (idx - 27) % 3;
        double q = 0.0;
        for (int i = 0; i <= c && c < D; i++) q += g.U[i][c] * F.MUc[d][i];
        F.Q[c][d] = q;
    } else if (idx < 42) {
        const int which = (idx - 36) / 3, i = (idx - 36) % 3;
        double acc = 0.0;
        if (i < D)
            for (int b = 0; b <= K; b++) acc += vmat<D>(F, b, K) * (which == 0 ? a.ht[b][i] : a.ut[b][i]);
        (which == 0 ? F.qt : F.uq)[i] = acc;
    }
}

// prep4: entry idx = l P + j of Kg J (needs prep1), the product every
// Hessian entry of fin_hess contracts with a column of J.
template <int D> __host__ __device__ inline void prep4(int idx, PairFin<D> &F)
{
    constexpr int NG = Dims<D>::NG, P = Dims<D>::P;
    const int l = idx / P, j = idx - (idx / P) * P;
    double kj = 0.0;
    for (int m = 0; m < NG; m++) kj += F.Kg[l][m] * F.J[m][j];
    F.KJ[l][j] = kj;
}

// `a` is caller-provided scratch (e.g. LDS) so that a thread does not have to
// keep the whole PairAlg in registers.
template <int D>
__host__ __device__ inline void pair_prepare(const PairGeom<D> &g, const double *pstar, const double *S1,
                                             const double *S2, int gn, PairFin<D> &F, PairAlg<D> &a)
{
    pair_alg<D>(g, pstar, a);
    for (int i = 0; i < PrepCount<D>::N1; i++) prep1<D>(g, a, S1, S2, gn, i, F);
    for (int i = 0; i < PrepCount<D>::N2; i++) prep2<D>(g, a, i, F);
    for (int i = 0; i < PrepCount<D>::N3; i++) prep3<D>(g, a, i, F);
    for (int i = 0; i < PrepCount<D>::N4; i++) prep4<D>(i, F);
}

template <int D>
__host__ __device__ inline void pair_prepare(const PairGeom<D> &g, const double *pstar, const double *S1,
                                             const double *S2, int gn, PairFin<D> &F)
{
    PairAlg<D> a;
    pair_prepare<D>(g, pstar, S1, S2, gn, F, a);
}

// Gradient entry i of one pair: (J^T v)_i.
template <int D> __host__ __device__ inline double fin_grad(const PairFin<D> &F, int i)
{
    double acc = 0.0;
    for (int l = 0; l < Dims<D>::NG; l++) acc += F.J[l][i] * F.v[l];
    return acc;
}

// Hessian entry (i, j) of one pair: (J^T Kg J)_ij + d2 Phi / d theta_i d theta_j.
template <int D> __host__ __device__ inline double fin_hess(const PairFin<D> &F, int i, int j)
{
    constexpr int NG = Dims<D>::NG;
    double acc = 0.0;
    for (int l = 0; l < NG; l++) acc += F.J[l][i] * F.KJ[l][j];
    if (i > j) {
        const int t = i;
        i = j;
        j = t;
    }
    const int r0 = 1 + D;  // first r slot
    if (i == 0) {
        if (j >= r0) acc += -F.va * F.rs3 * F.t[slot_row(D, j - r0)] * F.T[slot_col(D, j - r0)];
    } else if (i < r0) {
        const int c = i - 1;
        if (j < r0) {
            acc += 2.0 * F.VKK * F.Q[c][j - 1];
        } else {
            const int ii = slot_row(D, j - r0), jj = slot_col(D, j - r0);
            const double ec = (jj == c ? 1.0 : 0.0) - F.z[c] / F.s * F.T[jj];
            acc += -2.0 * (F.MUc[c][ii] * F.qt[jj] + F.uq[ii] * ec);
        }
    } else {
        const int i1 = slot_row(D, i - r0), j1 = slot_col(D, i - r0);
        const int i2 = slot_row(D, j - r0), j2 = slot_col(D, j - r0);
        const double dik = (i1 == i2) ? 1.0 : 0.0;
        acc += 2.0 * (dik * F.Hh[j2][j1] - (F.t[i1] * F.t[i2] * F.Hh[j2][j1] + F.t[i1] * F.T[j2] * F.Uh[i2][j1] +
                                          F.T[j1] * F.t[i2] * F.Uh[i1][j2] + F.T[j1] * F.T[j2] * F.Uu[i1][i2]) /
                                             F.s);
        acc += F.va * F.I * (3.0 * F.rs5 * F.t[i1] * F.T[j1] * F.t[i2] * F.T[j2] - F.rs3 * dik * F.T[j1] * F.T[j2]);
    }
    return acc;
}

}  // namespace gt
