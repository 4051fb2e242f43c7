// gt_ftplane.hip -- Fourier evaluation of Gaussian atoms on sample planes,
// and the complex L2 derivatives over hyperplane samples.
//
// Evaluation (FTGaussian3D.evaluate hyperplane / 2.5D branches,
// GaussDictCode/bin/FTGaussian3D.py:12-38, 113-148, 167-196, and
// FTGaussian2D.evaluate, FTGaussian2D.py:13-35, 53-100):
//     out[a][i0][i1] = sum_atoms I exp(-|L k|^2 / 2) exp(-i x . k),  |L k|^2 < cut
// with k = G_a p, p = (k0[i0], k1[i1]) and G_a the angle's 3x2 (or 2x2, 2x1)
// sample frame.  L is the FT-space shape: lower-triangular from the slots
// (rk = L k, rk_1 = r3 k_0 + r1 k_1, ...; 2D: rk_1 = r2 k_0 + r1 k_1); the
// isotropic forms are |k|^2 / r^2 (3D) and r^2 |k|^2 (2D).  cut = 30 (3D),
// 60 (2D).  For one atom and one angle the exponent is a 2x2 quadratic form
// in p and the phase is linear in p, so an FP64 setup per (atom, angle)
// reduces each sample to one quadratic, one compare, one exp2 and the
// hardware sine and cosine of a phase formed in turns from an exact split of
// the FP64 coefficients (as the grid kernel does; the FP64 evaluator uses
// sincospi).
//
// Derivatives (FTGaussian3D.derivs, :55-92, kernels :199-652), for every atom
// independently against all of C (the reference uses atom 0):
//     f = 1/2 sum |g - C|^2,  Df = Re sum (g - C) conj(dg),
//     DDf = Re sum dg conj(dg)^T + (g - C) conj(ddg),  g zero where |L k|^2 > 60.
// With E = exp(-|L k|^2/2 - i x.k), g = I E, every dg_j = c_j(p) E and
// ddg_jl = d_jl(p) E with c, d polynomials of degree <= 2 and <= 4 in p whose
// coefficients are real or purely imaginary (dg/dI = E, dg/dx_j = -i I K_j E,
// dg/dr_s = -I (M p)_row(s) (G p)_col(s) E, ...).  With w = (g - C) conj(E):
//     Df_j = Re sum w conj(c_j),  DDf_jl = Re sum |E|^2 c_j conj(c_l) + Re sum w conj(d_jl),
// so only Re of the even-degree and Im of the odd-degree moments of w, and
// the even moments of |E|^2, are needed: 24 FP32 moments per (atom, angle),
// accumulated column by column, then contracted in FP64 by the wave.  The
// isotropic 3D form is L = Id / r with the chain rule over the diagonal slots.
#include "gt_common.h"

namespace gt {

constexpr double KP_LOG2E = 1.4426950408889634;
constexpr double KP_TWO_PI = 6.283185307179586;

template <typename R> struct gt_atoms_t {
    int dim, n, iso;
    const R *I, *x, *r;
};

template <typename R> struct KpDev {
    int form, nA, m0, m1, sorted;
    const R *k0, *k1, *w0, *w1;
};

template <typename R> struct alignas(16) KpRec {
    R amp;
    R q[3];        // log2 exponent: q0 p0^2 + q1 p0 p1 + q2 p1^2
    R ch[2], cl[2];  // phase in turns per unit p (FP32: hi/lo pair; FP64: lo = 0)
    int lo[2], hi[2];
};

// Sample frame of angle a: k = G p (columns G[.][0], G[.][1]); D = 3 or 2.
template <typename R> __device__ inline void kp_frame(const KpDev<R> &P, int a, double G[3][2])
{
    for (int i = 0; i < 3; i++) G[i][0] = G[i][1] = 0.0;
    switch (P.form) {
    case GT_KP_3D:
        for (int i = 0; i < 3; i++) {
            G[i][0] = P.w0[(size_t)a * 3 + i];
            G[i][1] = P.w1[(size_t)a * 3 + i];
        }
        break;
    case GT_KP_2P5D:
        G[0][0] = P.w0[(size_t)a * 2];
        G[1][0] = P.w0[(size_t)a * 2 + 1];
        G[2][1] = 1.0;
        break;
    case GT_KP_2D_LINE:
        G[0][0] = P.w0[(size_t)a * 2];
        G[1][0] = P.w0[(size_t)a * 2 + 1];
        break;
    default:  // GT_KP_2D_GRID
        G[0][0] = 1.0;
        G[1][1] = 1.0;
        break;
    }
}

// FT-space shape L (rk = L k) of one atom: 3D slots (r0, r1, r2, r3 = L10,
// r4 = L21, r5 = L20); 2D slots (r0, r1, r2 = L10); iso: Id / r (3D), r Id (2D).
template <typename AR> __device__ inline void kp_shape(int D, int iso, const AR *r, double L[3][3])
{
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) L[i][k] = 0.0;
    if (iso) {
        const double v = (D == 3) ? 1.0 / (double)r[0] : (double)r[0];
        for (int i = 0; i < D; i++) L[i][i] = v;
    } else if (D == 3) {
        L[0][0] = r[0];
        L[1][1] = r[1];
        L[2][2] = r[2];
        L[1][0] = r[3];
        L[2][1] = r[4];
        L[2][0] = r[5];
    } else {
        L[0][0] = r[0];
        L[1][1] = r[1];
        L[1][0] = r[2];
    }
}

// One (atom, angle): A = (L G)^T (L G), phase coefficients G^T x / (2 pi), box
// of {p^T A p < cut}.
template <typename R, typename AR>
__device__ inline void kp_pair(int D, int iso, double I, const AR *x, const AR *r, const double G[3][2],
                               const KpDev<R> &P, double cut, KpRec<R> &rec)
{
    double L[3][3], M[3][2];
    kp_shape(D, iso, r, L);
    for (int i = 0; i < 3; i++)
        for (int c = 0; c < 2; c++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += L[i][k] * G[k][c];
            M[i][c] = s;
        }
    double A[2][2];
    for (int c = 0; c < 2; c++)
        for (int e = 0; e < 2; e++) {
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += M[i][c] * M[i][e];
            A[c][e] = s;
        }
    const double s2 = -0.5 * KP_LOG2E;
    rec.amp = (R)I;
    rec.q[0] = (R)(s2 * A[0][0]);
    rec.q[1] = (R)(2.0 * s2 * A[0][1]);
    rec.q[2] = (R)(s2 * A[1][1]);
    for (int c = 0; c < 2; c++) {
        double t = 0.0;
        for (int i = 0; i < D; i++) t += G[i][c] * (double)x[i];
        t /= KP_TWO_PI;
        rec.ch[c] = (R)t;
        rec.cl[c] = (R)(t - (double)rec.ch[c]);
    }
    // half-widths of the ellipse: sqrt(cut (A^-1)_cc); a singular A keeps the full range
    const double det = A[0][0] * A[1][1] - A[0][1] * A[0][1];
    const int m[2] = {P.m0, P.m1};
    const R *kv[2] = {P.k0, P.k1};
    const bool two = P.form != GT_KP_2D_LINE;
    for (int c = 0; c < 2; c++) {
        bool full = !P.sorted;
        double hw = 0.0;
        if (!two) {
            if (c == 1) {
                rec.lo[1] = 0;
                rec.hi[1] = 0;
                continue;
            }
            full = full || !(A[0][0] > 0.0);
            if (!full) hw = sqrt(cut / A[0][0]);
        } else {
            const double scale = fmax(A[0][0], A[1][1]);
            full = full || !(det > 1e-12 * scale * scale);
            if (!full) hw = sqrt(cut * A[1 - c][1 - c] / det);
        }
        if (full) {
            rec.lo[c] = 0;
            rec.hi[c] = m[c] - 1;
        } else {
            const double pad = 1e-5 * hw + 1e-7;
            rec.lo[c] = lower_idx(kv[c], m[c], -hw - pad);
            rec.hi[c] = upper_idx(kv[c], m[c], hw + pad);
        }
    }
}

template <typename R>
__global__ __launch_bounds__(256) void kp_setup_kernel(gt_atoms_t<R> A, KpDev<R> P, double cut, KpRec<R> *recs)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)A.n * P.nA) return;
    const int a = (int)(t / A.n), j = (int)(t - (long)a * A.n);
    const int D = A.dim;
    const int nr = A.iso ? 1 : D * (D + 1) / 2;
    double G[3][2];
    kp_frame(P, a, G);
    KpRec<R> rec;
    kp_pair(D, A.iso, (double)A.I[j], A.x + (size_t)j * D, A.r + (size_t)j * nr, G, P, cut, rec);
    recs[t] = rec;
}

// per-sample term, FP32: compensated phase as in the grid kernel
__device__ inline float kp_turns(float ch, float cl, float p, float &err)
{
    const float t = ch * p;
    const float e = fmaf(ch, p, -t);
    err += fmaf(cl, p, e);
    return t - rintf(t);
}

__device__ inline void kp_kahan(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

constexpr int KP_CHUNK = 128;

// One workgroup: 256 consecutive samples of one angle (flat index i0 * m1 + i1).
// Block b takes angle b % nA and the tile of centre-out rank b / nA: the tiles
// near k = 0 meet the most atoms and start first, the light outer ones fill
// the tail.
template <typename R>
__global__ __launch_bounds__(256) void kp_eval_kernel(const KpRec<R> *__restrict__ recs, int N, KpDev<R> P,
                                                      double cut, R *__restrict__ out)
{
    __shared__ KpRec<R> list[KP_CHUNK];
    __shared__ int wcnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long plane = (long)P.m0 * P.m1;
    const long tiles = (plane + 255) / 256;
    const int a = (int)(blockIdx.x % (unsigned)P.nA);
    const long base = (long)centre_out((int)(blockIdx.x / (unsigned)P.nA), (int)tiles) * 256;
    const long last = min(base + 255, plane - 1);
    const long idx = base + threadIdx.x;
    const bool valid = idx < plane;
    const int i0 = valid ? (int)(idx / P.m1) : 0, i1 = valid ? (int)(idx % P.m1) : 0;
    const int rlo = (int)(base / P.m1), rhi = (int)(last / P.m1);
    const int clo = (rlo == rhi) ? (int)(base % P.m1) : 0, chi = (rlo == rhi) ? (int)(last % P.m1) : P.m1 - 1;
    const R p0 = P.k0[i0];
    const R p1 = (P.form == GT_KP_2D_LINE) ? (R)0 : P.k1[i1];
    const R cutlog = (R)(-0.5 * KP_LOG2E * cut);
    const KpRec<R> *ra = recs + (size_t)a * N;
    R sr = 0, si = 0;
    float cr = 0.0f, ci = 0.0f;  // Kahan compensation (FP32 only)
    for (int b0 = 0; b0 < N; b0 += KP_CHUNK) {
        const int j = b0 + (int)threadIdx.x;
        bool hit = false;
        if (threadIdx.x < KP_CHUNK && j < N) {
            const KpRec<R> &r = ra[j];
            hit = r.lo[0] <= rhi && r.hi[0] >= rlo && r.lo[1] <= chi && r.hi[1] >= clo;
        }
        const unsigned long long msk = __ballot(hit);
        if (lane == 0) wcnt[wave] = __popcll(msk);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < 4; w++) {
            if (w < wave) off += wcnt[w];
            total += wcnt[w];
        }
        if (hit) list[off + __popcll(msk & ((1ull << lane) - 1ull))] = ra[j];
        __syncthreads();
        for (int e = 0; e < total; e++) {
            const KpRec<R> &r = list[e];
            if (r.lo[0] > i0 || r.hi[0] < i0 || r.lo[1] > i1 || r.hi[1] < i1) continue;
            const R ex = p0 * (r.q[0] * p0 + r.q[1] * p1) + r.q[2] * p1 * p1;
            if (!(ex > cutlog)) continue;
            if constexpr (sizeof(R) == 4) {
                float err = 0.0f;
                const float t = kp_turns(r.ch[0], r.cl[0], p0, err) + kp_turns(r.ch[1], r.cl[1], p1, err) + err;
                float sn, cs;
                ft_sincos_turns(t, sn, cs);
                const float am = r.amp * __builtin_amdgcn_exp2f(ex);
                kp_kahan(sr, cr, am * cs);
                kp_kahan(si, ci, -am * sn);
            } else {
                double t = r.ch[0] * p0 + r.ch[1] * p1;
                t -= rint(t);
                double sn, cs;
                sincospi(2.0 * t, &sn, &cs);
                const double am = r.amp * exp2(ex);
                sr += am * cs;
                si -= am * sn;
            }
        }
        __syncthreads();
    }
    if (valid) {
        out[2 * ((size_t)a * plane + idx)] = sr;
        out[2 * ((size_t)a * plane + idx) + 1] = si;
    }
}

// ---------------------------------------------------------------------------
// Derivatives
// ---------------------------------------------------------------------------
constexpr int KD_THREADS = 512;
constexpr int KD_UNR = 4;  // rows per step of the derivative walk (loads first)
constexpr int KD_WAVES = KD_THREADS / 64;
constexpr int KD_NV = 66;   // f part, 10 gradient entries, 55 upper Hessian entries
constexpr double KD_CUT = 60.0;

// Moment slots: Re w and |E|^2 over the even monomials, Im w over the odd ones.
// even: (0,0) (2,0) (1,1) (0,2) (4,0) (3,1) (2,2) (1,3) (0,4); odd: (1,0) (0,1) (3,0) (2,1) (1,2) (0,3)
__device__ inline int kd_even(int i, int j)
{
    const int d = i + j;
    return d == 0 ? 0 : (d == 2 ? 1 + j : 4 + j);
}
__device__ inline int kd_odd(int i, int j) { return (i + j) == 1 ? j : 2 + j; }

// Symmetric-tensor contractions of 2-vectors with the moment sets S(i, j) =
// sum w p0^i p1^j: T(u, v) etc. expanded over the monomials.
struct KdMom {
    double re[9], im[6], v[9];
};
__device__ inline double t2(const double *S, const double u[2], const double v[2])
{
    double s = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            const int n1 = a + b;
            s += u[a] * v[b] * S[kd_even(2 - n1, n1)];
        }
    return s;
}
__device__ inline double t1(const double *S, const double u[2]) { return u[0] * S[kd_odd(1, 0)] + u[1] * S[kd_odd(0, 1)]; }
__device__ inline double t3(const double *S, const double u[2], const double v[2], const double w[2])
{
    double s = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++) {
                const int n1 = a + b + c;
                s += u[a] * v[b] * w[c] * S[kd_odd(3 - n1, n1)];
            }
    return s;
}
__device__ inline double t4(const double *S, const double u[2], const double v[2], const double w[2], const double z[2])
{
    double s = 0.0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
            for (int c = 0; c < 2; c++)
                for (int d = 0; d < 2; d++) {
                    const int n1 = a + b + c + d;
                    s += u[a] * v[b] * w[c] * z[d] * S[kd_even(4 - n1, n1)];
                }
    return s;
}

// slot s of L -> row and column (rk_row = sum L[row][col] k_col), FTGaussian3D.py:231-234
__device__ inline int kd_row(int s) { return s < 3 ? s : (s == 3 ? 1 : 2); }
__device__ inline int kd_col(int s) { return s < 3 ? s : (s == 4 ? 1 : 0); }

// Entry e of [f part, Df (10), DDf upper (55)) of one (atom, angle), aniso
// parameters [I, x0..2, L slots 0..5).
__device__ inline double kd_entry(int e, double I, const double G[3][2], const double M[3][2], const KdMom &S)
{
    auto gv = [&](int j, double u[2]) { u[0] = G[j][0]; u[1] = G[j][1]; };
    auto mv = [&](int j, double u[2]) { u[0] = M[j][0]; u[1] = M[j][1]; };
    if (e == 0) return -0.5 * I * I * S.v[0] + I * S.re[0];
    if (e <= 10) {
        const int j = e - 1;
        if (j == 0) return S.re[0];
        double u[2], v[2];
        if (j <= 3) {
            gv(j - 1, u);
            return -I * t1(S.im, u);
        }
        mv(kd_row(j - 4), u);
        gv(kd_col(j - 4), v);
        return -I * t2(S.re, u, v);
    }
    // upper-triangle index -> (j, l)
    int k = e - 11, j = 0;
    while (k >= 10 - j) {
        k -= 10 - j;
        j++;
    }
    const int l = j + k;
    double u[2], v[2], w[2], z[2];
    if (j == 0) {
        if (l == 0) return S.v[0];
        if (l <= 3) {
            gv(l - 1, u);
            return -t1(S.im, u);
        }
        mv(kd_row(l - 4), u);
        gv(kd_col(l - 4), v);
        return -I * t2(S.v, u, v) - t2(S.re, u, v);
    }
    if (j <= 3) {
        gv(j - 1, u);
        if (l <= 3) {
            gv(l - 1, v);
            return I * I * t2(S.v, u, v) - I * t2(S.re, u, v);
        }
        mv(kd_row(l - 4), v);
        gv(kd_col(l - 4), w);
        return I * t3(S.im, u, v, w);
    }
    const int s = j - 4, t = l - 4;
    mv(kd_row(s), u);
    gv(kd_col(s), v);
    mv(kd_row(t), w);
    gv(kd_col(t), z);
    double val = I * I * t4(S.v, u, v, w, z) + I * t4(S.re, u, v, w, z);
    if (kd_row(s) == kd_row(t)) val -= I * t2(S.re, v, z);
    return val;
}

__device__ inline double kd_sum64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline float kd_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// One workgroup per atom; wave w takes angles w, w + KD_WAVES (8), ...  csum[a] =
// sum over angle a of |C|^2 (FP64).
// At most 128 VGPRs (amdgpu_waves_per_eu(4): two blocks of 8 waves per CU).
// The FP64 state that the row walk does not use (the atom's shape, the
// angle's frame and L G, the wave-summed moments) is kept in LDS rather than
// in registers across the walk, so that the walk runs without spills.
__global__ __launch_bounds__(KD_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void kd_kernel(gt_atoms A, KpDev<float> P, const float2 *__restrict__ C,
                                                        const double *__restrict__ csum, double *__restrict__ fo,
                                                        double *__restrict__ go, double *__restrict__ Ho)
{
    __shared__ double part[KD_WAVES][KD_NV];
    __shared__ double sx[3], srs[6], sL[3][3];                 // the atom: centre, shape slots, L
    __shared__ double sG[KD_WAVES][3][2], sM[KD_WAVES][3][2];  // each wave's angle: G, L G
    __shared__ KdMom sS[KD_WAVES];                             // each wave's moments (wave sums)
    const int j = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int iso = A.iso;
    const int nr = iso ? 1 : 6;
    const double I = A.I[j];
    if (threadIdx.x == 0) {
        double L[3][3];
        for (int i = 0; i < 3; i++) sx[i] = A.x[(size_t)j * 3 + i];
        kp_shape(3, iso, A.r + (size_t)j * nr, L);
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) sL[i][k] = L[i][k];
        // r as the aniso slots of L (iso: Id / r) for the pair setup
        srs[0] = L[0][0];
        srs[1] = L[1][1];
        srs[2] = L[2][2];
        srs[3] = L[1][0];
        srs[4] = L[2][1];
        srs[5] = L[2][0];
    }
    for (int e = lane; e < KD_NV; e += 64) part[wave][e] = 0.0;
    __syncthreads();
    const size_t plane = (size_t)P.m0 * P.m1;
    for (int a = wave; a < P.nA; a += KD_WAVES) {
        KpRec<float> rec;
        {
            double G[3][2];
            kp_frame(P, a, G);
            if (lane == 0)
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 2; c++) {
                        double s = 0.0;
                        for (int k = 0; k < 3; k++) s += sL[i][k] * G[k][c];
                        sM[wave][i][c] = s;
                        sG[wave][i][c] = G[i][c];
                    }
            kp_pair(3, 0, I, sx, srs, G, P, KD_CUT, rec);
        }
        const float cutlog = (float)(-0.5 * KP_LOG2E * KD_CUT);
        const float amp = (float)I;
        float Sre[9], Sim[6], Sv[9];
#pragma unroll
        for (int q = 0; q < 9; q++) Sre[q] = Sv[q] = 0.0f;
#pragma unroll
        for (int q = 0; q < 6; q++) Sim[q] = 0.0f;
        const float2 *Ca = C + (size_t)a * plane;
        for (int c = rec.lo[1] + lane; c <= rec.hi[1]; c += 64) {
            const float p1 = P.k1[c];
            float Are[5] = {0, 0, 0, 0, 0}, Aim[4] = {0, 0, 0, 0}, Av[5] = {0, 0, 0, 0, 0};
            float err1 = 0.0f;
            const float t1v = kp_turns(rec.ch[1], rec.cl[1], p1, err1);
            // rows in order, KD_UNR per step with their loads issued first;
            // outside the cut e = 0 and the sample adds exact zeros
            const float q2p1 = rec.q[2] * p1 * p1, q1p1 = rec.q[1] * p1;
            auto sample = [&](float p0, float2 cv) {
                const float ex = p0 * fmaf(rec.q[0], p0, q1p1) + q2p1;
                float err = err1;
                const float t = kp_turns(rec.ch[0], rec.cl[0], p0, err) + t1v + err;
                float sn, cs;
                ft_sincos_turns(t, sn, cs);
                const float e = ex >= cutlog ? __builtin_amdgcn_exp2f(ex) : 0.0f;
                // w = (I E - C) conj(E), conj(E) = e (cs + i sn)
                const float wr = amp * e * e - e * (cv.x * cs - cv.y * sn);
                const float wi = -e * (cv.x * sn + cv.y * cs);
                const float vv = e * e;
                float pw = 1.0f;
#pragma unroll
                for (int q = 0; q < 5; q++) {
                    Are[q] = fmaf(wr, pw, Are[q]);
                    Av[q] = fmaf(vv, pw, Av[q]);
                    if (q < 4) Aim[q] = fmaf(wi, pw, Aim[q]);
                    pw *= p0;
                }
            };
            int i = rec.lo[0];
            for (; i + KD_UNR - 1 <= rec.hi[0]; i += KD_UNR) {
                float p0v[KD_UNR];
                float2 cv[KD_UNR];
#pragma unroll
                for (int u = 0; u < KD_UNR; u++) {
                    p0v[u] = P.k0[i + u];
                    cv[u] = Ca[(size_t)(i + u) * P.m1 + c];
                }
#pragma unroll
                for (int u = 0; u < KD_UNR; u++) sample(p0v[u], cv[u]);
            }
            for (; i <= rec.hi[0]; i++) sample(P.k0[i], Ca[(size_t)i * P.m1 + c]);
            // combine the column sums with the powers of p1
            float pj[5] = {1.0f, p1, p1 * p1, p1 * p1 * p1, p1 * p1 * p1 * p1};
#pragma unroll
            for (int ii = 0; ii <= 4; ii++)
#pragma unroll
                for (int jj = 0; ii + jj <= 4; jj++) {
                    if ((ii + jj) % 2 == 0) {
                        const int s = (ii + jj) == 0 ? 0 : ((ii + jj) == 2 ? 1 + jj : 4 + jj);
                        Sre[s] = fmaf(Are[ii], pj[jj], Sre[s]);
                        Sv[s] = fmaf(Av[ii], pj[jj], Sv[s]);
                    } else {
                        const int s = (ii + jj) == 1 ? jj : 2 + jj;
                        Sim[s] = fmaf(Aim[ii], pj[jj], Sim[s]);
                    }
                }
        }
        KdMom &S = sS[wave];
#pragma unroll
        for (int q = 0; q < 9; q++) {
            const float re = kd_sum(Sre[q]), v = kd_sum(Sv[q]);
            if (lane == 0) {
                S.re[q] = (double)re;
                S.v[q] = (double)v;
            }
        }
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const float im = kd_sum(Sim[q]);
            if (lane == 0) S.im[q] = (double)im;
        }
        wave_sync();  // S, sG, sM written by lane 0, read by every lane
        for (int e = lane; e < KD_NV; e += 64) part[wave][e] += kd_entry(e, I, sG[wave], sM[wave], S);
        wave_sync();  // S, sG and sM are rewritten for the wave's next angle
    }
    __syncthreads();
    __shared__ double tot[KD_NV];
    if (threadIdx.x < KD_NV) {
        double s = 0.0;
        for (int w = 0; w < KD_WAVES; w++) s += part[w][threadIdx.x];
        if (threadIdx.x == 0) {
            double c2 = 0.0;
            for (int a = 0; a < P.nA; a++) c2 += csum[a];
            s += 0.5 * c2;
        }
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    // unpack [f, Df, DDf upper]; iso: chain rule over the diagonal slots, L_ii = 1 / r
    double g10[10], H10[10][10];
    for (int q = 0; q < 10; q++) g10[q] = tot[1 + q];
    int e = 11;
    for (int a = 0; a < 10; a++)
        for (int b = a; b < 10; b++) {
            H10[a][b] = H10[b][a] = tot[e];
            e++;
        }
    if (fo) fo[j] = tot[0];
    if (!iso) {
        for (int a = 0; a < 10; a++) {
            go[(size_t)j * 10 + a] = g10[a];
            if (Ho)
                for (int b = 0; b < 10; b++) Ho[(size_t)j * 100 + a * 10 + b] = H10[a][b];
        }
        return;
    }
    const double r = A.r[j];
    const double d1 = -1.0 / (r * r), d2 = 2.0 / (r * r * r);
    double g5[5], H5[5][5];
    for (int a = 0; a < 4; a++) g5[a] = g10[a];
    g5[4] = d1 * (g10[4] + g10[5] + g10[6]);
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 4; b++) H5[a][b] = H10[a][b];
    for (int a = 0; a < 4; a++) H5[a][4] = H5[4][a] = d1 * (H10[a][4] + H10[a][5] + H10[a][6]);
    double hrr = 0.0;
    for (int a = 4; a < 7; a++)
        for (int b = 4; b < 7; b++) hrr += H10[a][b];
    H5[4][4] = d1 * d1 * hrr + d2 * (g10[4] + g10[5] + g10[6]);
    for (int a = 0; a < 5; a++) {
        go[(size_t)j * 5 + a] = g5[a];
        if (Ho)
            for (int b = 0; b < 5; b++) Ho[(size_t)j * 25 + a * 5 + b] = H5[a][b];
    }
}

// csum[a] = sum over the samples of angle a of |C|^2, FP64, fixed order
__global__ __launch_bounds__(256) void kd_csum_kernel(const float2 *__restrict__ C, size_t plane, double *csum)
{
    __shared__ double part[4];
    const size_t a = blockIdx.x;
    double s = 0.0;
    for (size_t i = threadIdx.x; i < plane; i += 256) {
        const float2 v = C[a * plane + i];
        s += (double)v.x * v.x + (double)v.y * v.y;
    }
    s = kd_sum64(s);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) csum[a] = part[0] + part[1] + part[2] + part[3];
}

template <typename R> __global__ void kp_fill_kernel(R *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

template <typename R, typename KP> int kp_check(const KP *P, int dim, const char *who)
{
    GT_CHECK_ARG(P != nullptr, "%s: planes is NULL", who);
    GT_CHECK_ARG(P->form >= GT_KP_3D && P->form <= GT_KP_2D_GRID, "%s: unknown sample form %d", who, P->form);
    const int pdim = (P->form <= GT_KP_2P5D) ? 3 : 2;
    GT_CHECK_ARG(pdim == dim, "%s: sample form of dimension %d for atoms of dimension %d", who, pdim, dim);
    GT_CHECK_ARG(P->nA >= 1 && P->m0 >= 1 && P->m1 >= 1, "%s: nA, m0, m1 must be >= 1", who);
    GT_CHECK_ARG(P->form != GT_KP_2D_LINE || P->m1 == 1, "%s: the 2D line form has m1 == 1", who);
    GT_CHECK_ARG(P->form != GT_KP_2D_GRID || P->nA == 1, "%s: the 2D grid form has nA == 1", who);
    GT_CHECK_ARG((long)P->nA * P->m0 * P->m1 < (1L << 40), "%s: too many samples", who);
    GT_CHECK_ARG(P->k0 != nullptr && (P->form == GT_KP_2D_LINE || P->k1 != nullptr), "%s: k0 / k1 missing", who);
    GT_CHECK_ARG(P->form == GT_KP_2D_GRID || P->w0 != nullptr, "%s: w0 missing", who);
    GT_CHECK_ARG(P->form != GT_KP_3D || P->w1 != nullptr, "%s: w1 missing", who);
    GT_CHECK_ARG(P->sorted == 0 || P->sorted == 1, "%s: sorted must be 0 or 1", who);
    return 0;
}

template <typename R, typename KP> KpDev<R> kp_dev(const KP *P)
{
    KpDev<R> d;
    d.form = P->form;
    d.nA = P->nA;
    d.m0 = P->m0;
    d.m1 = P->m1;
    d.sorted = P->sorted;
    d.k0 = P->k0;
    d.k1 = P->k1;
    d.w0 = P->w0;
    d.w1 = P->w1;
    return d;
}

template <typename R, typename AT, typename KP>
int kp_eval(const AT *atoms, const KP *planes, R *out, void *ws, size_t ws_bytes, void *stream, const char *who)
{
    int rc;
    GT_CHECK_ARG(atoms != nullptr, "%s: atoms is NULL", who);
    GT_CHECK_ARG(atoms->dim == 2 || atoms->dim == 3, "%s: atoms.dim must be 2 or 3", who);
    GT_CHECK_ARG(atoms->n >= 0 && (atoms->iso == 0 || atoms->iso == 1), "%s: bad atom count or iso flag", who);
    GT_CHECK_ARG(atoms->n == 0 || (atoms->I && atoms->x && atoms->r), "%s: atom arrays missing", who);
    if ((rc = kp_check<R>(planes, atoms->dim, who))) return rc;
    GT_CHECK_ARG(out != nullptr, "%s: out is NULL", who);
    const size_t need = align_up((size_t)atoms->n * planes->nA * sizeof(KpRec<R>));
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("%s: workspace of %zu bytes given, %zu needed", who, ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "%s: workspace must be 64-byte aligned", who);
    hipStream_t st = (hipStream_t)stream;
    const KpDev<R> P = kp_dev<R>(planes);
    const size_t nout = 2 * (size_t)planes->nA * planes->m0 * planes->m1;
    if (atoms->n == 0) {
        hipLaunchKernelGGL(kp_fill_kernel<R>, dim3((unsigned)std::min<size_t>((nout + 255) / 256, 65535)), dim3(256), 0,
                           st, out, nout);
        GT_HIP(hipGetLastError());
        return 0;
    }
    const double cut = atoms->dim == 3 ? 30.0 : 60.0;
    KpRec<R> *recs = (KpRec<R> *)ws;
    const long npair = (long)atoms->n * planes->nA;
    GT_CHECK_ARG(npair < (1L << 31) * 256L, "%s: too many (atom, angle) pairs", who);
    gt_atoms_t<R> A{atoms->dim, atoms->n, atoms->iso, atoms->I, atoms->x, atoms->r};
    hipLaunchKernelGGL(kp_setup_kernel<R>, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, st, A, P, cut, recs);
    const long tiles = ((long)planes->m0 * planes->m1 + 255) / 256;
    const long blocks = tiles * planes->nA;
    GT_CHECK_ARG(blocks < (1L << 31), "%s: too many samples", who);
    prof_begin(st, GT_K_FT);
    hipLaunchKernelGGL(kp_eval_kernel<R>, dim3((unsigned)blocks), dim3(256), 0, st, recs, atoms->n, P, cut, out);
    prof_end(st, GT_K_FT);
    GT_HIP(hipGetLastError());
    return 0;
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_ft_planes_workspace(int n_atoms, const gt_kplanes *planes)
{
    if (!planes || n_atoms < 0) return 0;
    return align_up((size_t)n_atoms * planes->nA * sizeof(KpRec<float>));
}

extern "C" int gt_ft_planes_eval(const gt_atoms *atoms, const gt_kplanes *planes, float *out, void *ws, size_t ws_bytes,
                                 void *stream)
{
    return kp_eval<float>(atoms, planes, out, ws, ws_bytes, stream, "gt_ft_planes_eval");
}

extern "C" size_t gt_ft_planes_workspace_f64(int n_atoms, const gt_kplanes64 *planes)
{
    if (!planes || n_atoms < 0) return 0;
    return align_up((size_t)n_atoms * planes->nA * sizeof(KpRec<double>));
}

extern "C" int gt_ft_planes_eval_f64(const gt_atoms64 *atoms, const gt_kplanes64 *planes, double *out, void *ws,
                                     size_t ws_bytes, void *stream)
{
    return kp_eval<double>(atoms, planes, out, ws, ws_bytes, stream, "gt_ft_planes_eval_f64");
}

extern "C" size_t gt_ft_derivs_workspace(const gt_atoms *atoms, const gt_kplanes *planes)
{
    if (!atoms || !planes) return 0;
    return align_up((size_t)planes->nA * sizeof(double));
}

extern "C" int gt_ft_derivs(const gt_atoms *atoms, const gt_kplanes *planes, const float *C, double *f, double *g,
                            double *H, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms))) return rc;
    GT_CHECK_ARG(atoms->dim == 3, "gt_ft_derivs: 3D atoms only (the reference has no 2D derivatives)");
    if ((rc = kp_check<float>(planes, 3, "gt_ft_derivs"))) return rc;
    GT_CHECK_ARG(C != nullptr, "gt_ft_derivs: C is NULL");
    GT_CHECK_ARG(g != nullptr || atoms->n == 0, "gt_ft_derivs: g is NULL");
    const size_t need = gt_ft_derivs_workspace(atoms, planes);
    if (need > ws_bytes || ws == nullptr) {
        set_error("gt_ft_derivs: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    if (atoms->n == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    const KpDev<float> P = kp_dev<float>(planes);
    double *csum = (double *)ws;
    const size_t plane = (size_t)planes->m0 * planes->m1;
    hipLaunchKernelGGL(kd_csum_kernel, dim3(planes->nA), dim3(256), 0, st, (const float2 *)C, plane, csum);
    prof_begin(st, GT_K_FT);
    hipLaunchKernelGGL(kd_kernel, dim3(atoms->n), dim3(KD_THREADS), 0, st, *atoms, P, (const float2 *)C, csum, f, g, H);
    prof_end(st, GT_K_FT);
    GT_HIP(hipGetLastError());
    return 0;
}
