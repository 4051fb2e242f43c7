// gt_psf.hip -- the point-spread ("single particle") projector: detector
// values integrated over each pixel, volume values integrated over each voxel.
//
// SingleParticleGaussTomo binds GaussDictCode/bin/point_spread_functions.py:
//   RadProj          :371-385   P[a][k] = int_{pixel k} R_a(p) dp
//   L2derivs_RadProj :1077-1115 f = 1/2 sum_k (P_k - C_k)^2, its gradient and Hessian
//   derivs_RadProj   :1314-1350 f = sum_k P_k C_k, its gradient and Hessian
//   VolProj          :168-197   u[v] = int_{voxel v} f(y) dy
// with R the line integral of one atom (exp(-30 m) / |UT|, cut where m > 1)
// and f its volume profile.  A pixel is the square of side |p[1] - p[0]|
// around each detector coordinate (per axis; a voxel likewise).
//
// The reference samples each pixel on an 11 x 11 grid that starts one step
// inside the pixel's lower edge and ends one step beyond its upper edge
// (:436-455, 497-520), returns the atom's total for pixels that contain its
// bounding circle and 0 gradients for them (:1012-1015), and uses the first
// axis' spacing for both axes (:563-566).  Here the pixel integral itself is
// computed: for one (atom, angle) the detector profile is alpha 2^(q(d)) with
// q the record's quadratic form (d = p - p*), and the pixel is intersected with
// the bounding box of the support {m <= 1} and integrated on pieces no longer
// than two marginal standard deviations by Gauss-Legendre rules of 3 to 8
// points sized to the piece (error below about 1e-7 of peak x width), the
// second detector axis in closed form when its interval is long (psf_pixel).
//
// Derivatives use the projector's moment contraction (pair_finish) with the
// point moments replaced by pixel-integrated ones, M_k[mono] = int_k mono(d)
// E(d) dd: the linear terms (gradient, and the residual-weighted second
// derivatives) are sums over pixels of weighted M_k; the Gauss-Newton term
// sum_k dP_k dP_k^T, a product of two pixel integrals, enters through the
// products sum_k M_k[m] M_k[m'] of the degree <= 2 monomials.
//
// Rule GT_PSF_SAMPLED reproduces the reference's own numbers instead
// (__rf2 / __rf3 :413-527, __drf2 / __drf3 :924-1075, __if2 / __if3
// :222-291): per (atom, cell) with n the distance of the (projected) centre
// from the cell centre and R = 1 / min diag U, n + R < s gives the atom's
// total (I (pi/30) / det U, or I sqrt(pi/30) / det U in 2D) with no
// derivative, n - R > s gives 0, and otherwise the integrand (no cut-off) is
// sampled at lo + k (hi - lo) / 10, k = 1..11, per axis on [lo, hi] = the
// cell cut to [-R, R] about the centre, each sample weighted (hi - lo) / 11;
// s is half the first detector axis' spacing for both detector axes and half
// the second volume axis' spacing for every volume axis.  The derivatives are
// those of the integrand at the samples (the same moment contraction, with the
// samples for nodes).
#include "gt_common.h"

namespace gt {

// Gauss-Legendre rules of 3, 4, 5, 6 and 8 points on [-1, 1], concatenated
// (offsets 0, 3, 7, 12, 18).
__device__ constexpr float GLX[26] = {
    -0.7745966692414834f, 0.0000000000000000f, 0.7745966692414834f, -0.8611363115940526f, -0.3399810435848563f,
    0.3399810435848563f, 0.8611363115940526f, -0.9061798459386640f, -0.5384693101056831f, 0.0000000000000000f,
    0.5384693101056831f, 0.9061798459386640f, -0.9324695142031521f, -0.6612093864662645f, -0.2386191860831969f,
    0.2386191860831969f, 0.6612093864662645f, 0.9324695142031521f, -0.9602898564975362f, -0.7966664774136267f,
    -0.5255324099163290f, -0.1834346424956498f, 0.1834346424956498f, 0.5255324099163290f, 0.7966664774136267f,
    0.9602898564975362f};
__device__ constexpr float GLW[26] = {
    0.5555555555555557f, 0.8888888888888888f, 0.5555555555555557f, 0.3478548451374537f, 0.6521451548625462f,
    0.6521451548625462f, 0.3478548451374537f, 0.2369268850561894f, 0.4786286704993662f, 0.5688888888888890f,
    0.4786286704993662f, 0.2369268850561894f, 0.1713244923791697f, 0.3607615730481389f, 0.4679139345726914f,
    0.4679139345726914f, 0.3607615730481389f, 0.1713244923791697f, 0.1012285362903767f, 0.2223810344533743f,
    0.3137066458778870f, 0.3626837833783618f, 0.3626837833783618f, 0.3137066458778870f, 0.2223810344533743f,
    0.1012285362903767f};

// Points of the rule for an interval of length `len` over which the
// integrand varies on the scale `sc` (a standard deviation): the error of
// each is below about 1e-7 of (peak x len) wherever the interval sits on a
// Gaussian (3 points to len <= 0.5 sc, 4 to 1.0, 5 to 1.4, 6 to 2.0; 8 beyond).
__device__ inline int psf_rule(float len, float sc, int &off)
{
    const float H = len / sc;
    if (H <= 0.5f) { off = 0; return 3; }
    if (H <= 1.0f) { off = 3; return 4; }
    if (H <= 1.4f) { off = 7; return 5; }
    if (H <= 2.0f) { off = 12; return 6; }
    off = 18;
    return 8;
}

// Conditional standard deviation along an axis whose log2-exponent
// coefficient is qd (< 0): the integrand's scale along that axis.
__device__ inline float psf_csig(float qd) { return qd < 0.0f ? rsqrtf(-1.3862943611198906f * qd) : 3.0e38f; }

constexpr int PSF_MAXPIECES = 8;
constexpr int PSF_MS = 40;  // floats per (atom, angle) moment record: S1[15] | Q[21] | f part | pad

struct alignas(16) PsfRec {
    PairRec r;           // alpha, p*, q, inclusive pixel box of the pixels meeting the support
    float sl[2], sh[2];  // support bounding box in d = p - p*
    float sig[2];        // marginal standard deviation of exp(-30 m) per detector axis
    float s[2];          // pixel half-widths
};
static_assert(sizeof(PsfRec) == 96, "PsfRec must stay 96 bytes");

// R = 1 / min diag U (:25-30) and the atom's total I (pi/30) / det U, or
// I sqrt(pi/30) / det U in 2D (:60-65): the constants of the sampled rule.
template <int D> __device__ inline void psf_rule_consts(const PairGeom<D> &g, double &rad, double &total)
{
    double dmin = g.U[0][0], dprod = 1.0;
    for (int i = 0; i < D; i++) {
        dmin = fmin(dmin, g.U[i][i]);
        dprod *= g.U[i][i];
    }
    rad = 1.0 / dmin;
    const double c = 3.14159265358979323846 / 30.0;
    total = g.I * (D == 3 ? c : sqrt(c)) / dprod;
}

// Record of the sampled rule: sl = (R, total), s = (s, s) with s half the
// first detector axis' spacing; pixel box = the pixels within |R| + s of p*
// per axis (every other pixel has n - R > s, or n >= |R| + s > s - R, so the
// rule gives it 0).
template <int D>
__device__ inline void psf_setup_sampled(const PairGeom<D> &g, const float *const *det, const int *nd, float hs0,
                                         PsfRec &R)
{
    constexpr int K = Dims<D>::K;
    pair_setup<D>(g, det, nd, 0, R.r);
    double rad, total;
    psf_rule_consts<D>(g, rad, total);
    const double pstar[2] = {(double)R.r.c_hi[0] + (double)R.r.c_lo[0], (double)R.r.c_hi[1] + (double)R.r.c_lo[1]};
    R.sl[0] = (float)rad;
    R.sl[1] = (float)total;
    for (int k = 0; k < 2; k++) {
        R.sh[k] = 0.0f;
        R.sig[k] = 0.0f;
        R.s[k] = hs0;
        if (k >= K) {
            R.r.lo[k] = 0;
            R.r.hi[k] = 0;
            continue;
        }
        const double reach = fabs(rad) + (double)hs0;
        const double pad = 1e-5 * reach + 1e-7;
        R.r.lo[k] = lower_idx(det[k], nd[k], pstar[k] - reach - pad);
        R.r.hi[k] = upper_idx(det[k], nd[k], pstar[k] + reach + pad);
    }
}

template <int D>
__device__ inline void psf_setup(const PairGeom<D> &g, const float *const *det, const int *nd, const float *hs,
                                 PsfRec &R)
{
    constexpr int K = Dims<D>::K;
    pair_setup<D>(g, det, nd, 0, R.r);
    const double pstar[2] = {(double)R.r.c_hi[0] + (double)R.r.c_lo[0], (double)R.r.c_hi[1] + (double)R.r.c_lo[1]};
    PairAlg<D> a;
    pair_alg<D>(g, pstar, a);
    double off[2] = {0.0, 0.0}, hw[2] = {0.0, 0.0}, sg[2] = {1.0, 1.0};
    bool empty = false;
    if (K == 2) {
        const double A00 = a.B[0][0], A01 = a.B[0][1], A11 = a.B[1][1];
        const double det2 = A00 * A11 - A01 * A01;
        if (!(det2 > 1e-12 * A00 * A11 && A00 > 0.0 && A11 > 0.0)) {
            empty = true;
        } else {
            const double i00 = A11 / det2, i01 = -A01 / det2, i11 = A00 / det2;
            const double b0 = a.B[0][2], b1 = a.B[1][2];
            off[0] = -(i00 * b0 + i01 * b1);
            off[1] = -(i01 * b0 + i11 * b1);
            const double rho2 = 1.0 - a.B[2][2] + (b0 * (i00 * b0 + i01 * b1) + b1 * (i01 * b0 + i11 * b1));
            if (!(rho2 >= 0.0)) empty = true;
            hw[0] = sqrt(fmax(rho2, 0.0) * i00);
            hw[1] = sqrt(fmax(rho2, 0.0) * i11);
            sg[0] = sqrt(i00 / 60.0);
            sg[1] = sqrt(i11 / 60.0);
        }
    } else {
        const double A = a.B[0][0], b = a.B[0][1];
        if (!(A > 0.0)) {
            empty = true;
        } else {
            off[0] = -b / A;
            const double rho2 = 1.0 - a.B[1][1] + b * b / A;
            if (!(rho2 >= 0.0)) empty = true;
            hw[0] = sqrt(fmax(rho2, 0.0) / A);
            sg[0] = sqrt(1.0 / (60.0 * A));
        }
    }
    for (int k = 0; k < 2; k++) {
        R.sl[k] = (float)(off[k] - hw[k]);
        R.sh[k] = (float)(off[k] + hw[k]);
        R.sig[k] = (float)sg[k];
        R.s[k] = (k < K) ? hs[k] : 0.0f;
        if (k >= K) {
            R.r.lo[k] = 0;
            R.r.hi[k] = 0;
            continue;
        }
        if (empty) {
            R.r.lo[k] = 1;
            R.r.hi[k] = 0;
            continue;
        }
        const double pad = 1e-5 * (hw[k] + hs[k]) + 1e-7;
        R.r.lo[k] = lower_idx(det[k], nd[k], pstar[k] + off[k] - hw[k] - hs[k] - pad);
        R.r.hi[k] = upper_idx(det[k], nd[k], pstar[k] + off[k] + hw[k] + hs[k] + pad);
    }
}

// half width of the cells of an ascending coordinate vector, from its first
// spacing (the reference's assumption of a uniform grid, :533, 563)
__device__ inline float psf_cell_half(const float *v) { return 0.5f * fabsf(v[1] - v[0]); }

template <int D>
__global__ __launch_bounds__(256) void psf_setup_kernel(gt_atoms A, gt_geom G, PsfRec *recs, int rule)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)A.n * G.nA) return;
    const int a = (int)(idx / A.n), j = (int)(idx - (long)a * A.n);
    PairGeom<D> g;
    load_atom<D>(A, j, g);
    load_angle<D>(G, a, g);
    const float *det[2] = {G.p0, G.p1};
    const int nd[2] = {G.n0, G.n1};
    const float hs[2] = {psf_cell_half(G.p0), D == 3 ? psf_cell_half(G.p1) : 0.0f};
    PsfRec R;
    if (rule == GT_PSF_SAMPLED)
        psf_setup_sampled<D>(g, det, nd, hs[0], R);
    else
        psf_setup<D>(g, det, nd, hs, R);
    recs[idx] = R;
}

// Pieces and step of the composite rule on [L, U]: pieces of at most 2 sigma.
__device__ inline int psf_pieces(float L, float U, float sig, float &h)
{
    const int n = min(PSF_MAXPIECES, max(1, (int)ceilf((U - L) / (2.0f * sig))));
    h = (U - L) / (float)n;
    return n;
}

// Inner integrals in closed form: G[i] = int_l^h u^i exp(-k u^2) du, i < NJ,
// from erf / erfc (G0) and the recursion
//   G_i = [-u^(i-1) exp(-k u^2) / (2k)]_l^h + (i - 1) / (2k) G_(i-2).
// Used only when sqrt(k) (h - l) >= sqrt(2), where the erf difference is well
// conditioned; erfc of the side away from the peak avoids cancellation in
// the tails.
template <int NJ> __device__ inline void psf_inner(float k, float l, float h, float *G)
{
    const float sk = sqrtf(k);
    const float al = sk * l, ah = sk * h;
    float F;
    if (al >= 0.0f)
        F = erfcf(al) - erfcf(ah);
    else if (ah <= 0.0f)
        F = erfcf(-ah) - erfcf(-al);
    else
        F = erff(ah) - erff(al);
    G[0] = 0.88622692545275801f / sk * F;  // sqrt(pi) / 2
    if (NJ > 1) {
        const float el = expf(-al * al), eh = expf(-ah * ah);
        const float i2k = 0.5f / k;
        G[1] = (el - eh) * i2k;
        if (NJ > 2) G[2] = fmaf(l * el - h * eh, i2k, G[0] * i2k);
        if (NJ > 3) G[3] = fmaf(l * l * el - h * h * eh, i2k, 2.0f * G[1] * i2k);
        if (NJ > 4) G[4] = fmaf(l * l * l * el - h * h * h * eh, i2k, 3.0f * G[2] * i2k);
    }
}

// int t^j exp(-k (t - mu)^2) over the same interval, j < NJ, from the G's (t = mu + u).
template <int NJ> __device__ inline void psf_shift(float mu, const float *G, float *T)
{
    T[0] = G[0];
    if (NJ > 1) T[1] = fmaf(mu, G[0], G[1]);
    if (NJ > 2) T[2] = fmaf(mu, fmaf(mu, G[0], 2.0f * G[1]), G[2]);
    if (NJ > 3) T[3] = fmaf(mu, fmaf(mu, fmaf(mu, G[0], 3.0f * G[1]), 3.0f * G[2]), G[3]);
    if (NJ > 4) T[4] = fmaf(mu, fmaf(mu, fmaf(mu, fmaf(mu, G[0], 4.0f * G[1]), 6.0f * G[2]), 4.0f * G[3]), G[4]);
}

template <int NM> struct PsfDeg {
    static constexpr int NJ = NM == 1 ? 1 : NM <= 3 ? 2 : NM <= 6 ? 3 : NM <= 10 ? 4 : 5;
};

// NP x NP Gauss-Legendre rule on [L0, U0] x [L1, U1] (NP = 3 or 4), moments as psf_pixel.
template <int NP, int NM>
__device__ inline void psf_tensor(const float *q, float L0, float U0, float L1, float U1, float *M)
{
    constexpr float X3[3] = {-0.7745966692414834f, 0.0f, 0.7745966692414834f};
    constexpr float W3[3] = {0.5555555555555557f, 0.8888888888888888f, 0.5555555555555557f};
    constexpr float X4[4] = {-0.8611363115940526f, -0.3399810435848563f, 0.3399810435848563f, 0.8611363115940526f};
    constexpr float W4[4] = {0.3478548451374537f, 0.6521451548625462f, 0.6521451548625462f, 0.3478548451374537f};
    const float kap = (float)KAPPA;
    const float m0 = 0.5f * (L0 + U0), r0 = 0.5f * (U0 - L0);
    const float m1 = 0.5f * (L1 + U1), r1 = 0.5f * (U1 - L1);
    constexpr int NJ = PsfDeg<NM>::NJ;
#pragma unroll
    for (int a = 0; a < NP; a++) {
        const float d0 = fmaf(r0, NP == 3 ? X3[a] : X4[a], m0);
        const float wa = r0 * r1 * (NP == 3 ? W3[a] : W4[a]);
        // B_j = sum over the inner nodes of w d1^j, then M[i, j] += d0^i B_j
        float B[NJ];
#pragma unroll
        for (int j = 0; j < NJ; j++) B[j] = 0.0f;
#pragma unroll
        for (int b = 0; b < NP; b++) {
            const float d1 = fmaf(r1, NP == 3 ? X3[b] : X4[b], m1);
            const float e = pair_q(q, d0, d1);
            float w = (e >= kap) ? wa * (NP == 3 ? W3[b] : W4[b]) * __builtin_amdgcn_exp2f(e) : 0.0f;
#pragma unroll
            for (int j = 0; j < NJ; j++) {
                B[j] += w;
                w *= d1;
            }
        }
        if (NM == 1) {
            M[0] += B[0];
            continue;
        }
        const float P0[5] = {1.0f, d0, d0 * d0, d0 * d0 * d0, d0 * d0 * d0 * d0};
#pragma unroll
        for (int deg = 0; deg <= 4; deg++)
#pragma unroll
            for (int j = 0; j <= deg; j++) {
                const int m = deg * (deg + 1) / 2 + j;
                if (m < NM) M[m] = fmaf(P0[deg - j], B[j < NJ ? j : 0], M[m]);
            }
    }
}



// Pixel-integrated moments of one (atom, angle) over the pixel centred at
// (pc0, pc1): M[mono] = int mono(d) 2^q(d) dd (NM = 1: the value only).
// The pixel is cut to the support's bounding box.  An axis whose interval
// is longer than two marginal standard deviations is integrated in closed
// form (the profile is a 1D Gaussian along it: the last axis for 2D
// detectors, given the first-axis node); otherwise by composite 8-point
// Gauss-Legendre on pieces of at most 2 sigma.  The closed form integrates
// the untruncated Gaussian, which differs from the profile cut at m > 1 by
// less than exp(-30) of its peak.  The number of points per piece follows
// the piece's length over the integrand's scale along the axis (psf_rule).
template <int K, int NM>
__device__ inline void psf_pixel(const PsfRec &R, float pc0, float pc1, float *M)
{
    // 2D detectors: NM counts monomials in (d0, d1); 1D: powers of d0
    constexpr int NJ = (K == 1) ? NM : PsfDeg<NM>::NJ;
    constexpr float LN2 = 0.69314718055994531f;
#pragma unroll
    for (int m = 0; m < NM; m++) M[m] = 0.0f;
    const float c0 = (pc0 - R.r.c_hi[0]) - R.r.c_lo[0];
    const float L0 = fmaxf(c0 - R.s[0], R.sl[0]), U0 = fminf(c0 + R.s[0], R.sh[0]);
    if (!(L0 < U0)) return;
    float h0;
    const int n0 = psf_pieces(L0, U0, R.sig[0], h0);
    const float *q = R.r.q;
    const float kap = (float)KAPPA;
    if (K == 1) {
        if (n0 > 1) {
            // exponent (natural log) -k (d - mu)^2 + peak
            const float k = -LN2 * q[0];
            const float mu = LN2 * q[3] / (2.0f * k);
            const float peak = fmaf(fmaf(q[0], mu, q[3]), mu, q[5]);
            if (peak < kap) return;
            float G[NJ], T[NJ];
            psf_inner<NJ>(k, L0 - mu, U0 - mu, G);
            psf_shift<NJ>(mu, G, T);
            const float w = __builtin_amdgcn_exp2f(peak);
#pragma unroll
            for (int m = 0; m < NM; m++) M[m] = w * T[m];
            return;
        }
        int o0;
        const int g0 = psf_rule(h0, psf_csig(q[0]), o0);
        for (int i0 = 0; i0 < n0; i0++) {
            const float mid = L0 + h0 * ((float)i0 + 0.5f);
            for (int a = 0; a < g0; a++) {
                const float d = fmaf(0.5f * h0, GLX[o0 + a], mid);
                const float e = fmaf(fmaf(q[0], d, q[3]), d, q[5]);
                if (e >= kap) {
                    float w = 0.5f * h0 * GLW[o0 + a] * __builtin_amdgcn_exp2f(e);
#pragma unroll
                    for (int m = 0; m < NM; m++) {
                        M[m] = fmaf(w, 1.0f, M[m]);
                        w *= d;
                    }
                }
            }
        }
        return;
    }
    const float c1 = (pc1 - R.r.c_hi[1]) - R.r.c_lo[1];
    const float L1 = fmaxf(c1 - R.s[1], R.sl[1]), U1 = fminf(c1 + R.s[1], R.sh[1]);
    if (!(L1 < U1)) return;
    float h1;
    const int n1 = psf_pieces(L1, U1, R.sig[1], h1);
    const bool closed = n1 > 1;
    const float k1 = -LN2 * q[2];
    int o0, o1;
    const int g0 = psf_rule(h0, psf_csig(q[0]), o0);
    const int g1 = psf_rule(h1, psf_csig(q[2]), o1);
    if (!closed && n0 == 1 && g0 <= 4 && g1 <= 4) {
        // the common case of an atom wider than the pixel: one tensor rule,
        // unrolled, so that the lanes of a wave (neighbouring pixels of the
        // same atom) run the same instructions
        if (g0 <= 3 && g1 <= 3)
            psf_tensor<3, NM>(q, L0, U0, L1, U1, M);
        else
            psf_tensor<4, NM>(q, L0, U0, L1, U1, M);
        return;
    }
    for (int i0 = 0; i0 < n0; i0++) {
        const float mid0 = L0 + h0 * ((float)i0 + 0.5f);
        for (int a = 0; a < g0; a++) {
            const float d0 = fmaf(0.5f * h0, GLX[o0 + a], mid0);
            const float wa = 0.5f * h0 * GLW[o0 + a];
            float P0[5] = {1.0f, d0, d0 * d0, d0 * d0 * d0, d0 * d0 * d0 * d0};
            if (closed) {
                // given d0 the profile is exp(-k1 (d1 - mu)^2) 2^peak
                const float mu = LN2 * fmaf(q[1], d0, q[4]) / (2.0f * k1);
                const float peak = pair_q(q, d0, mu);
                if (peak < kap) continue;
                float G[NJ], T[NJ];
                psf_inner<NJ>(k1, L1 - mu, U1 - mu, G);
                psf_shift<NJ>(mu, G, T);
                const float w = wa * __builtin_amdgcn_exp2f(peak);
#pragma unroll
                for (int deg = 0; deg <= 4; deg++)
#pragma unroll
                    for (int j = 0; j <= deg; j++) {
                        const int m = deg * (deg + 1) / 2 + j;
                        if (m < NM) M[m] = fmaf(w * P0[deg - j], T[j < NJ ? j : 0], M[m]);
                    }
                continue;
            }
            for (int i1 = 0; i1 < n1; i1++) {
                const float mid1 = L1 + h1 * ((float)i1 + 0.5f);
                for (int b = 0; b < g1; b++) {
                    const float d1 = fmaf(0.5f * h1, GLX[o1 + b], mid1);
                    const float e = pair_q(q, d0, d1);
                    if (e >= kap) {
                        const float w = wa * (0.5f * h1 * GLW[o1 + b]) * __builtin_amdgcn_exp2f(e);
                        if (NM == 1) {
                            M[0] += w;
                        } else {
                            // monomials by degree: 1 | d0 d1 | d0^2 d0d1 d1^2 | ... (mono2 order)
                            float p1[5] = {1.0f, d1, d1 * d1, d1 * d1 * d1, d1 * d1 * d1 * d1};
#pragma unroll
                            for (int deg = 0; deg <= 4; deg++)
#pragma unroll
                                for (int j = 0; j <= deg; j++) {
                                    const int m = deg * (deg + 1) / 2 + j;
                                    if (m < NM) M[m] = fmaf(w, P0[deg - j] * p1[j], M[m]);
                                }
                        }
                    }
                }
            }
        }
    }
}

// The sampled rule for one (atom, angle) and the pixel centred at (pc0, pc1):
// returns the pixel's value; M[mono] = w sum over the samples of mono(d)
// 2^q(d) (zero for total and zero pixels, whose value has no derivative).
// Offsets d = p - p*: the pixel centre sits at c = pc - p* (the reference's
// <x - y, W> is -c), the samples of an axis at L + k h, k = 1..11, with L =
// max(-R, c - s), h = (min(R, c + s) - L) / 10, each of weight 10 h / 11.
template <int K, int NM>
__device__ inline float psf_pixel_sampled(const PsfRec &R, float pc0, float pc1, float *M)
{
    constexpr int NJ = (K == 1) ? NM : PsfDeg<NM>::NJ;
#pragma unroll
    for (int m = 0; m < NM; m++) M[m] = 0.0f;
    const float c0 = (pc0 - R.r.c_hi[0]) - R.r.c_lo[0];
    const float c1 = (K == 2) ? (pc1 - R.r.c_hi[1]) - R.r.c_lo[1] : 0.0f;
    const float rad = R.sl[0], s = R.s[0];
    const float n = (K == 2) ? sqrtf(fmaf(c0, c0, c1 * c1)) : fabsf(c0);
    if (n + rad < s) return R.sl[1];
    if (n - rad > s) return 0.0f;
    const float *q = R.r.q;
    const float L0 = fmaxf(-rad, c0 - s), h0 = 0.1f * (fminf(rad, c0 + s) - L0);
    float w = (10.0f / 11.0f) * h0;
    if (K == 1) {
#pragma unroll
        for (int k = 1; k <= 11; k++) {
            const float d = fmaf((float)k, h0, L0);
            float e = __builtin_amdgcn_exp2f(fmaf(fmaf(q[0], d, q[3]), d, q[5]));
#pragma unroll
            for (int m = 0; m < NM; m++) {
                M[m] += e;
                e *= d;
            }
        }
    } else {
        const float L1 = fmaxf(-rad, c1 - s), h1 = 0.1f * (fminf(rad, c1 + s) - L1);
        w *= (10.0f / 11.0f) * h1;
#pragma unroll 1
        for (int k = 1; k <= 11; k++) {
            const float d0 = fmaf((float)k, h0, L0);
            // exponent along d1 for this d0: q2 d1^2 + (q1 d0 + q4) d1 + (q0 d0^2 + q3 d0 + q5)
            const float b = fmaf(q[1], d0, q[4]), c = fmaf(fmaf(q[0], d0, q[3]), d0, q[5]);
            float B[NJ];
#pragma unroll
            for (int j = 0; j < NJ; j++) B[j] = 0.0f;
#pragma unroll
            for (int l = 1; l <= 11; l++) {
                const float d1 = fmaf((float)l, h1, L1);
                float e = __builtin_amdgcn_exp2f(fmaf(fmaf(q[2], d1, b), d1, c));
#pragma unroll
                for (int j = 0; j < NJ; j++) {
                    B[j] += e;
                    e *= d1;
                }
            }
            if (NM == 1) {
                M[0] += B[0];
                continue;
            }
            const float P0[5] = {1.0f, d0, d0 * d0, d0 * d0 * d0, d0 * d0 * d0 * d0};
#pragma unroll
            for (int deg = 0; deg <= 4; deg++)
#pragma unroll
                for (int j = 0; j <= deg; j++) {
                    const int m = deg * (deg + 1) / 2 + j;
                    if (m < NM) M[m] = fmaf(P0[deg - j], B[j < NJ ? j : 0], M[m]);
                }
        }
    }
#pragma unroll
    for (int m = 0; m < NM; m++) M[m] *= w;
    return R.r.alpha * M[0];
}

__device__ inline void psf_kahan(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

constexpr int PSF_CHUNK = 64;

// Stage the records of one angle whose pixel box meets [rlo, rhi] x [clo, chi], in atom order.
__device__ inline int psf_stage(const PsfRec *__restrict__ ra, int base, int N, int rlo, int rhi, int clo, int chi,
                                PsfRec *list, int *wcnt)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = base + (int)threadIdx.x;
    bool hit = false;
    if (threadIdx.x < PSF_CHUNK && j < N) {
        const PsfRec &r = ra[j];
        hit = r.r.lo[0] <= rhi && r.r.hi[0] >= rlo && r.r.lo[1] <= chi && r.r.hi[1] >= clo;
    }
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        if (w < wave) off += wcnt[w];
        total += wcnt[w];
    }
    if (hit) list[off + __popcll(m & ((1ull << lane) - 1ull))] = ra[j];
    __syncthreads();
    return total;
}

// Forward: block = (angle, 16 x 16 pixels) for 2D detectors, numbered centre
// tiles first (block b: angle b % nA, tile of Morton rank b / nA over the
// centre-out tile ranks, 2^e0 x 2^e1), or (angle = blockIdx.y, 256 pixels) for
// 1D ones.
template <int K, int RULE>
__global__ __launch_bounds__(256) void psf_forward_kernel(const PsfRec *__restrict__ recs, int N, int nA, int n0,
                                                          int n1, const float *__restrict__ p0,
                                                          const float *__restrict__ p1, float *__restrict__ R, int e0,
                                                          int e1)
{
    __shared__ PsfRec list[PSF_CHUNK];
    __shared__ int wcnt[4];
    int a = blockIdx.y;
    int i0, i1, rlo, rhi, clo, chi;
    if (K == 2) {
        const int ntr = (n0 + 15) / 16, ntc = (n1 + 15) / 16;
        a = (int)(blockIdx.x % (unsigned)nA);
        const int e[3] = {e0, e1, 0};
        int t[3];
        morton_ranks(blockIdx.x / (unsigned)nA, e, t);
        if (t[0] >= ntr || t[1] >= ntc) return;  // whole block: no tile of this rank
        const int tr = centre_out(t[0], ntr), tcc = centre_out(t[1], ntc);
        i0 = tr * 16 + (int)(threadIdx.x >> 4);
        i1 = tcc * 16 + (int)(threadIdx.x & 15);
        rlo = tr * 16;
        rhi = min(rlo + 15, n0 - 1);
        clo = tcc * 16;
        chi = min(clo + 15, n1 - 1);
    } else {
        i0 = blockIdx.x * 256 + (int)threadIdx.x;
        i1 = 0;
        rlo = blockIdx.x * 256;
        rhi = min(rlo + 255, n0 - 1);
        clo = chi = 0;
    }
    const bool valid = i0 < n0 && i1 < n1;
    const float pc0 = p0[min(i0, n0 - 1)];
    const float pc1 = (K == 2) ? p1[min(i1, n1 - 1)] : 0.0f;
    const PsfRec *ra = recs + (size_t)a * N;
    float s = 0.0f, c = 0.0f;
    for (int base = 0; base < N; base += PSF_CHUNK) {
        const int total = psf_stage(ra, base, N, rlo, rhi, clo, chi, list, wcnt);
        for (int e = 0; e < total; e++) {
            const PsfRec &r = list[e];
            if (!valid || r.r.lo[0] > i0 || r.r.hi[0] < i0 || r.r.lo[1] > i1 || r.r.hi[1] < i1) continue;
            float M[1], v;
            if (RULE == GT_PSF_SAMPLED) {
                v = psf_pixel_sampled<K, 1>(r, pc0, pc1, M);
            } else {
                psf_pixel<K, 1>(r, pc0, pc1, M);
                v = r.r.alpha * M[0];
            }
            psf_kahan(s, c, v);
        }
        __syncthreads();
    }
    if (valid) R[((size_t)a * n0 + i0) * n1 + i1] = s;
}

// Moments: one wave per (atom, angle) pair over the pixels of its box.
template <int K, int MODE, int RULE>
__global__ __launch_bounds__(256) void psf_moments_kernel(const PsfRec *__restrict__ recs, int N, int nA, int n0,
                                                          int n1, const float *__restrict__ p0,
                                                          const float *__restrict__ p1, const float *__restrict__ C,
                                                          int order, float *__restrict__ mom)
{
    constexpr int NM = (K == 2) ? 15 : 5, NL = (K == 2) ? 6 : 3, NQ = NL * (NL + 1) / 2;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long pair = (long)blockIdx.x * 4 + wave;
    if (pair >= (long)N * nA) return;
    const PsfRec R = recs[pair];
    const int a = (int)(pair / N);
    float S1[NM], Q[NQ], fp = 0.0f;
#pragma unroll
    for (int m = 0; m < NM; m++) S1[m] = 0.0f;
#pragma unroll
    for (int m = 0; m < NQ; m++) Q[m] = 0.0f;
    const int rows = R.r.hi[0] - R.r.lo[0] + 1, cols = R.r.hi[1] - R.r.lo[1] + 1;
    const long count = (rows > 0 && cols > 0) ? (long)rows * cols : 0;
    const float *Ca = C + (size_t)a * n0 * n1;
    for (long t = lane; t < count; t += 64) {
        const int i0 = R.r.lo[0] + (int)(t / cols), i1 = R.r.lo[1] + (int)(t % cols);
        float M[NM], P;
        if (RULE == GT_PSF_SAMPLED) {
            P = psf_pixel_sampled<K, NM>(R, p0[i0], (K == 2) ? p1[i1] : 0.0f, M);
        } else {
            psf_pixel<K, NM>(R, p0[i0], (K == 2) ? p1[i1] : 0.0f, M);
            P = R.r.alpha * M[0];
        }
        const float cv = Ca[(size_t)i0 * n1 + i1];
        if (MODE == GT_MODE_L2) {
            const float res = P - cv;
#pragma unroll
            for (int m = 0; m < NM; m++) S1[m] = fmaf(res, M[m], S1[m]);
            fp = fmaf(P, 0.5f * P - cv, fp);
            if (order > 1) {
                int k = 0;
#pragma unroll
                for (int u = 0; u < NL; u++)
#pragma unroll
                    for (int v = u; v < NL; v++) {
                        Q[k] = fmaf(M[u], M[v], Q[k]);
                        k++;
                    }
            }
        } else {
#pragma unroll
            for (int m = 0; m < NM; m++) S1[m] = fmaf(cv, M[m], S1[m]);
            fp = fmaf(P, cv, fp);
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
        for (int m = 0; m < NM; m++) S1[m] += __shfl_xor(S1[m], o);
#pragma unroll
        for (int m = 0; m < NQ; m++) Q[m] += __shfl_xor(Q[m], o);
        fp += __shfl_xor(fp, o);
    }
    if (lane == 0) {
        float *out = mom + pair * PSF_MS;
        for (int m = 0; m < NM; m++) out[m] = S1[m];
        for (int m = NM; m < 15; m++) out[m] = 0.0f;
        for (int m = 0; m < NQ; m++) out[15 + m] = Q[m];
        for (int m = NQ; m < 21; m++) out[15 + m] = 0.0f;
        out[36] = fp;
        out[37] = out[38] = out[39] = 0.0f;
    }
}

// FP64 sum of C^2 per angle
__global__ __launch_bounds__(256) void psf_sumsq_kernel(const float *__restrict__ C, long per, double *__restrict__ out)
{
    __shared__ double red[4];
    const int a = blockIdx.x;
    double s = 0.0;
    for (long i = threadIdx.x; i < per; i += 256) {
        const double v = C[(size_t)a * per + i];
        s += v * v;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[a] = red[0] + red[1] + red[2] + red[3];
}

// Contraction and reduction over angles: one workgroup per atom, lane = angle.
template <int D, int MODE>
__global__ __launch_bounds__(64) void psf_finish_kernel(gt_atoms A, gt_geom G, const PsfRec *__restrict__ recs,
                                                        const float *__restrict__ mom,
                                                        const double *__restrict__ sumC2, int want_hess,
                                                        double *__restrict__ fo, double *__restrict__ go,
                                                        double *__restrict__ Ho)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, P = Dm::P, NM = Dm::NM, NL = (K == 2) ? 6 : 3, NH = P * (P + 1) / 2, NV = 1 + P + NH;
    // row of lane l: [F, grad (P), upper Hessian by rows (NH)]; the lane
    // accumulates its pairs' Hessian entries there directly (a P x P array
    // per lane in registers would spill)
    __shared__ double red[64][NV + 1];
    const int j = blockIdx.x, lane = threadIdx.x;
    const int N = A.n;
    double grad[P], F = 0.0;
    for (int i = 0; i < P; i++) grad[i] = 0.0;
    double *row = red[lane];
    for (int k = 0; k < NV; k++) row[k] = 0.0;
    PairGeom<D> g;
    load_atom<D>(A, j, g);
    for (int a = lane; a < G.nA; a += 64) {
        const size_t pi = (size_t)a * N + j;
        const PsfRec R = recs[pi];
        const float *m = mom + pi * PSF_MS;
        if (MODE == GT_MODE_L2) F += 0.5 * sumC2[a];
        F += (double)m[36];
        if (R.r.lo[0] > R.r.hi[0] || (K == 2 && R.r.lo[1] > R.r.hi[1])) continue;
        load_angle<D>(G, a, g);
        double S1[NM], Q[NL * NL];
        for (int k = 0; k < NM; k++) S1[k] = m[k];
        {
            int k = 0;
            for (int u = 0; u < NL; u++)
                for (int v = u; v < NL; v++) {
                    Q[u * NL + v] = Q[v * NL + u] = m[15 + k];
                    k++;
                }
        }
        const double pstar[2] = {(double)R.r.c_hi[0] + (double)R.r.c_lo[0], (double)R.r.c_hi[1] + (double)R.r.c_lo[1]};
        pair_finish_to<D>(g, pstar, S1, nullptr, MODE == GT_MODE_L2, want_hess, grad, HessUpper<P>{row + 1 + P},
                          MODE == GT_MODE_L2 ? Q : nullptr);
    }
    row[0] = F;
    for (int i = 0; i < P; i++) row[1 + i] = grad[i];
    __syncthreads();
    for (int col = lane; col < NV; col += 64) {
        double s = 0.0;
        for (int i = 0; i < 64; i++) s += red[i][col];
        red[0][col] = s;
    }
    __syncthreads();
    const double *v = red[0];
    if (!A.iso) {
        // the lanes write the gradient and the full Hessian, one entry each
        if (lane == 0 && fo) fo[j] = v[0];
        if (lane < P) go[(size_t)j * P + lane] = v[1 + lane];
        if (Ho && want_hess)
            for (int e = lane; e < P * P; e += 64) {
                const int i = e / P, l = e - (e / P) * P;
                const int lo = i < l ? i : l, hi = i < l ? l : i;
                Ho[(size_t)j * P * P + e] = v[1 + P + lo * P - lo * (lo - 1) / 2 + (hi - lo)];
            }
        return;
    }
    if (lane != 0) return;
    if (fo) fo[j] = v[0];
    double Hf[P][P];
    {
        int k = 1 + P;
        for (int i = 0; i < P; i++)
            for (int l = i; l < P; l++) {
                Hf[i][l] = Hf[l][i] = v[k];
                k++;
            }
    }
    // U = rho Id: the diagonal slots move together
    constexpr int PI = 2 + D;
    auto src = [](int l, int &b, int &n) {
        if (l <= D) {
            b = l;
            n = 1;
        } else {
            b = 1 + D;
            n = D;
        }
    };
    for (int l = 0; l < PI; l++) {
        int b, n;
        src(l, b, n);
        double s = 0.0;
        for (int u = 0; u < n; u++) s += v[1 + b + u];
        go[(size_t)j * PI + l] = s;
    }
    if (Ho && want_hess)
        for (int l = 0; l < PI; l++)
            for (int m2 = 0; m2 < PI; m2++) {
                int b0, c0, b1, c1;
                src(l, b0, c0);
                src(m2, b1, c1);
                double s = 0.0;
                for (int u = 0; u < c0; u++)
                    for (int w = 0; w < c1; w++) s += Hf[b0 + u][b1 + w];
                Ho[((size_t)j * PI + l) * PI + m2] = s;
            }
}

// ---------------------------------------------------------------------------
// Voxel-integrated volume render
// ---------------------------------------------------------------------------
struct alignas(16) PsfVolRec {
    float amp;           // I sqrt(30 / pi)
    float xh[3], xl[3];  // centre as float pairs
    float u[6];          // U slots scaled by sqrt(30 log2 e): exponent = -|U'(y - x)|^2
    float hw[3];         // support half-widths (rows of U^-1)
    int lo[3], hi[3];    // voxel box of the voxels meeting the support
    float pad[2];
};
static_assert(sizeof(PsfVolRec) == 96, "PsfVolRec must stay 96 bytes");

constexpr double PSF_SQRT_NEG_KAPPA = 6.578818071009935;  // sqrt(30 log2 e)

// Rule GT_PSF_SAMPLED: hw = (R, total, s) with s half the second axis'
// spacing, voxel box = the voxels within |R| + s of x per axis.
template <int D>
__global__ __launch_bounds__(256) void psf_vol_setup_kernel(gt_atoms A, gt_vol V, PsfVolRec *recs, int rule)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    constexpr int NR = Dims<D>::NR;
    PairGeom<D> g;
    load_atom<D>(A, j, g);
    PsfVolRec R;
    R.amp = (float)(g.I * sqrt(30.0 / 3.14159265358979323846));
    for (int i = 0; i < 3; i++) {
        const double xv = (i < D) ? g.x[i] : 0.0;
        R.xh[i] = (float)xv;
        R.xl[i] = (float)(xv - (double)R.xh[i]);
    }
    for (int k = 0; k < 6; k++) R.u[k] = 0.0f;
    for (int k = 0; k < NR; k++) R.u[k] = (float)(PSF_SQRT_NEG_KAPPA * g.U[slot_row(D, k)][slot_col(D, k)]);
    R.pad[0] = R.pad[1] = 0.0f;
    const float *y[3] = {V.y0, V.y1, V.y2};
    const int m[3] = {V.m0, V.m1, V.m2};
    if (rule == GT_PSF_SAMPLED) {
        double rad, total;
        psf_rule_consts<D>(g, rad, total);
        const float sv = psf_cell_half(V.y1);
        R.hw[0] = (float)rad;
        R.hw[1] = (float)total;
        R.hw[2] = sv;
        for (int c = 0; c < 3; c++) {
            if (c >= D) {
                R.lo[c] = 0;
                R.hi[c] = 0;
                continue;
            }
            const double reach = fabs(rad) + (double)sv;
            const double pad = 1e-5 * reach + 1e-7;
            R.lo[c] = lower_idx(y[c], m[c], g.x[c] - reach - pad);
            R.hi[c] = upper_idx(y[c], m[c], g.x[c] + reach + pad);
        }
        recs[j] = R;
        return;
    }
    const float hv[3] = {psf_cell_half(V.y0), psf_cell_half(V.y1), D == 3 ? psf_cell_half(V.y2) : 0.0f};
    double Ui[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool ok = true;
    for (int i = 0; i < D; i++) ok = ok && g.U[i][i] != 0.0;
    if (ok)
        for (int c = D - 1; c >= 0; c--)
            for (int col = 0; col < D; col++) {
                double s = (c == col) ? 1.0 : 0.0;
                for (int k = c + 1; k < D; k++) s -= g.U[c][k] * Ui[k][col];
                Ui[c][col] = s / g.U[c][c];
            }
    for (int c = 0; c < 3; c++) {
        R.hw[c] = 0.0f;
        if (c >= D) {
            R.lo[c] = 0;
            R.hi[c] = 0;
            continue;
        }
        if (!ok) {
            R.lo[c] = 1;
            R.hi[c] = 0;
            continue;
        }
        double hw = 0.0;
        for (int k = 0; k < D; k++) hw += Ui[c][k] * Ui[c][k];
        hw = sqrt(hw);
        R.hw[c] = (float)hw;
        const double pad = 1e-5 * (hw + hv[c]) + 1e-7;
        R.lo[c] = lower_idx(y[c], m[c], g.x[c] - hw - hv[c] - pad);
        R.hi[c] = upper_idx(y[c], m[c], g.x[c] + hw + hv[c] + pad);
    }
    recs[j] = R;
}

// exponent of one node: -|U'(y - x)|^2 with y - x from the float pairs
template <int D> __device__ inline float psf_vexp(const PsfVolRec &R, const float *dy)
{
    if (D == 3) {
        const float a0 = fmaf(R.u[0], dy[0], fmaf(R.u[3], dy[1], R.u[5] * dy[2]));
        const float a1 = fmaf(R.u[1], dy[1], R.u[4] * dy[2]);
        const float a2 = R.u[2] * dy[2];
        return -fmaf(a0, a0, fmaf(a1, a1, a2 * a2));
    }
    const float a0 = fmaf(R.u[0], dy[0], R.u[2] * dy[1]);
    const float a1 = R.u[1] * dy[1];
    return -fmaf(a0, a0, a1 * a1);
}

// NP^D tensor Gauss-Legendre rule over the box [L, L + 2r] (NP = 3 or 4).
template <int D, int NP> __device__ inline float psf_vtensor(const PsfVolRec &R, const float *L, const float *r)
{
    constexpr float X3[3] = {-0.7745966692414834f, 0.0f, 0.7745966692414834f};
    constexpr float W3[3] = {0.5555555555555557f, 0.8888888888888888f, 0.5555555555555557f};
    constexpr float X4[4] = {-0.8611363115940526f, -0.3399810435848563f, 0.3399810435848563f, 0.8611363115940526f};
    constexpr float W4[4] = {0.3478548451374537f, 0.6521451548625462f, 0.6521451548625462f, 0.3478548451374537f};
    const float kap = (float)KAPPA;
    float acc = 0.0f, dy[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int a = 0; a < NP; a++) {
        dy[0] = fmaf(r[0], (NP == 3 ? X3[a] : X4[a]) + 1.0f, L[0]);
        const float wa = NP == 3 ? W3[a] : W4[a];
#pragma unroll
        for (int b = 0; b < NP; b++) {
            dy[1] = fmaf(r[1], (NP == 3 ? X3[b] : X4[b]) + 1.0f, L[1]);
            const float wb = wa * (NP == 3 ? W3[b] : W4[b]);
            if (D == 2) {
                const float e = psf_vexp<D>(R, dy);
                acc += (e >= kap) ? wb * __builtin_amdgcn_exp2f(e) : 0.0f;
                continue;
            }
#pragma unroll
            for (int c = 0; c < NP; c++) {
                dy[2] = fmaf(r[2], (NP == 3 ? X3[c] : X4[c]) + 1.0f, L[2]);
                const float e = psf_vexp<D>(R, dy);
                acc += (e >= kap) ? wb * (NP == 3 ? W3[c] : W4[c]) * __builtin_amdgcn_exp2f(e) : 0.0f;
            }
        }
    }
    return acc * r[0] * r[1] * (D == 3 ? r[2] : 1.0f);
}

// voxel integral of one atom: composite Gauss-Legendre over voxel x support
// box, pieces of at most 2 marginal standard deviations, points per piece
// from the piece's length over the profile's scale along the axis
// (psf_rule; the scale is the conditional standard deviation, 1 / sqrt(2 ln2
// |column of U'|^2)).
template <int D> __device__ inline float psf_voxel(const PsfVolRec &R, const float *yc, const float *hv)
{
    float L[3], h[3];
    int n[3] = {1, 1, 1}, g[3] = {1, 1, 1}, o[3] = {0, 0, 0};
    const float cn[3] = {R.u[0] * R.u[0],
                         D == 3 ? fmaf(R.u[3], R.u[3], R.u[1] * R.u[1]) : fmaf(R.u[2], R.u[2], R.u[1] * R.u[1]),
                         D == 3 ? fmaf(R.u[5], R.u[5], fmaf(R.u[4], R.u[4], R.u[2] * R.u[2])) : 0.0f};
    bool one = true;
    int gmax = 0;
    for (int c = 0; c < D; c++) {
        const float dc = (yc[c] - R.xh[c]) - R.xl[c];
        L[c] = fmaxf(dc - hv[c], -R.hw[c]);
        const float U = fminf(dc + hv[c], R.hw[c]);
        if (!(L[c] < U)) return 0.0f;
        n[c] = psf_pieces(L[c], U, R.hw[c] * 0.12909944487358055f, h[c]);  // sigma = hw / sqrt(60)
        g[c] = psf_rule(h[c], psf_csig(-cn[c]), o[c]);
        one = one && n[c] == 1;
        gmax = max(gmax, g[c]);
    }
    if (one && gmax <= 4) {
        const float r[3] = {0.5f * h[0], 0.5f * h[1], 0.5f * h[2]};
        return R.amp * (gmax <= 3 ? psf_vtensor<D, 3>(R, L, r) : psf_vtensor<D, 4>(R, L, r));
    }
    const float kap = (float)KAPPA;
    float acc = 0.0f;
    float dy[3] = {0.0f, 0.0f, 0.0f};
    for (int i0 = 0; i0 < n[0]; i0++)
        for (int a = 0; a < g[0]; a++) {
            dy[0] = fmaf(0.5f * h[0], GLX[o[0] + a], L[0] + h[0] * ((float)i0 + 0.5f));
            const float w0 = 0.5f * h[0] * GLW[o[0] + a];
            for (int i1 = 0; i1 < n[1]; i1++)
                for (int b = 0; b < g[1]; b++) {
                    dy[1] = fmaf(0.5f * h[1], GLX[o[1] + b], L[1] + h[1] * ((float)i1 + 0.5f));
                    const float w1 = w0 * 0.5f * h[1] * GLW[o[1] + b];
                    if (D == 2) {
                        const float e = psf_vexp<D>(R, dy);
                        if (e >= kap) acc = fmaf(w1, __builtin_amdgcn_exp2f(e), acc);
                        continue;
                    }
                    for (int i2 = 0; i2 < n[2]; i2++)
                        for (int c = 0; c < g[2]; c++) {
                            dy[2] = fmaf(0.5f * h[2], GLX[o[2] + c], L[2] + h[2] * ((float)i2 + 0.5f));
                            const float e = psf_vexp<D>(R, dy);
                            if (e >= kap) acc = fmaf(w1 * 0.5f * h[2] * GLW[o[2] + c], __builtin_amdgcn_exp2f(e), acc);
                        }
                }
        }
    return R.amp * acc;
}

// The sampled rule for one atom and the voxel centred at yc: n = |yc - x|,
// n + R < s gives the atom's total, n - R > s gives 0, otherwise amp times
// the samples' sum of exp(-30 |U d|^2) at d = L + k h per axis (k = 1..11, L =
// max(-R, c - s), h = (min(R, c + s) - L) / 10, c = yc - x), each of weight
// prod 10 h / 11.
template <int D> __device__ inline float psf_voxel_sampled(const PsfVolRec &R, const float *yc)
{
    float c[3] = {0.0f, 0.0f, 0.0f}, L[3] = {0.0f, 0.0f, 0.0f}, h[3] = {0.0f, 0.0f, 0.0f};
    float n2 = 0.0f;
    for (int i = 0; i < D; i++) {
        c[i] = (yc[i] - R.xh[i]) - R.xl[i];
        n2 = fmaf(c[i], c[i], n2);
    }
    const float rad = R.hw[0], s = R.hw[2], n = sqrtf(n2);
    if (n + rad < s) return R.hw[1];
    if (n - rad > s) return 0.0f;
    float w = R.amp;
    for (int i = 0; i < D; i++) {
        L[i] = fmaxf(-rad, c[i] - s);
        h[i] = 0.1f * (fminf(rad, c[i] + s) - L[i]);
        w *= (10.0f / 11.0f) * h[i];
    }
    float acc = 0.0f, dy[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int a = 1; a <= 11; a++) {
        dy[0] = fmaf((float)a, h[0], L[0]);
        if (D == 2) {
#pragma unroll
            for (int b = 1; b <= 11; b++) {
                dy[1] = fmaf((float)b, h[1], L[1]);
                acc += __builtin_amdgcn_exp2f(psf_vexp<D>(R, dy));
            }
            continue;
        }
#pragma unroll 1
        for (int b = 1; b <= 11; b++) {
            dy[1] = fmaf((float)b, h[1], L[1]);
#pragma unroll
            for (int k = 1; k <= 11; k++) {
                dy[2] = fmaf((float)k, h[2], L[2]);
                acc += __builtin_amdgcn_exp2f(psf_vexp<D>(R, dy));
            }
        }
    }
    return w * acc;
}

// Block = brick of 4 x 8 x 8 voxels (3D) or 16 x 16 (2D), thread = voxel,
// PV_GROUPS groups of 256 threads on the same brick: group g takes every
// PV_GROUPS-th staged record, and the groups' compensated sums are added in
// group order at the end (deterministic).  Blocks are numbered bricks nearest
// the volume centre first (Morton order over centre-out brick ranks): the
// central bricks meet most atoms, and with several times the blocks that fit
// at once the light outer bricks fill the tail.
constexpr int PV_GROUPS = 4, PV_THREADS = PV_GROUPS * 256, PV_CHUNK = 256;

template <int D, int RULE>
__global__ __launch_bounds__(PV_THREADS) void psf_volume_kernel(const PsfVolRec *__restrict__ recs, int N, gt_vol V,
                                                                float *__restrict__ u, int e0, int e1, int e2)
{
    __shared__ PsfVolRec list[PV_CHUNK];
    __shared__ int wcnt[PV_CHUNK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = threadIdx.x >> 8, t = threadIdx.x & 255;
    const int B0 = (D == 3) ? 4 : 16, B1 = (D == 3) ? 8 : 16, B2 = (D == 3) ? 8 : 1;
    const int nb0 = (V.m0 + B0 - 1) / B0, nb1 = (V.m1 + B1 - 1) / B1, nb2 = (V.m2 + B2 - 1) / B2;
    const int e[3] = {e0, e1, e2};
    int tb[3];
    morton_ranks(blockIdx.x, e, tb);
    if (tb[0] >= nb0 || tb[1] >= nb1 || tb[2] >= nb2) return;  // whole block
    const int bb0 = centre_out(tb[0], nb0), bb1 = centre_out(tb[1], nb1), bb2 = centre_out(tb[2], nb2);
    const int i0 = bb0 * B0 + t / (B1 * B2), i1 = bb1 * B1 + (t / B2) % B1, i2 = bb2 * B2 + t % B2;
    const int blo[3] = {bb0 * B0, bb1 * B1, bb2 * B2};
    const int bhi[3] = {min(blo[0] + B0 - 1, V.m0 - 1), min(blo[1] + B1 - 1, V.m1 - 1), min(blo[2] + B2 - 1, V.m2 - 1)};
    const bool valid = i0 < V.m0 && i1 < V.m1 && i2 < V.m2;
    const float yc[3] = {V.y0[min(i0, V.m0 - 1)], V.y1[min(i1, V.m1 - 1)], (D == 3) ? V.y2[min(i2, V.m2 - 1)] : 0.0f};
    const float hv[3] = {psf_cell_half(V.y0), psf_cell_half(V.y1), D == 3 ? psf_cell_half(V.y2) : 0.0f};
    float s = 0.0f, c = 0.0f;
    for (int base = 0; base < N; base += PV_CHUNK) {
        // stage the chunk's records that meet the brick (waves 0..3 scan, atom order kept)
        const int j = base + (int)threadIdx.x;
        bool hit = false;
        if (threadIdx.x < PV_CHUNK && j < N) {
            const PsfVolRec &r = recs[j];
            hit = r.lo[0] <= bhi[0] && r.hi[0] >= blo[0] && r.lo[1] <= bhi[1] && r.hi[1] >= blo[1] &&
                  r.lo[2] <= bhi[2] && r.hi[2] >= blo[2];
        }
        const unsigned long long msk = __ballot(hit);
        if (wave < PV_CHUNK / 64 && lane == 0) wcnt[wave] = __popcll(msk);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < PV_CHUNK / 64; w++) {
            if (w < wave) off += wcnt[w];
            total += wcnt[w];
        }
        if (hit) list[off + __popcll(msk & ((1ull << lane) - 1ull))] = recs[j];
        __syncthreads();
        for (int e = grp; e < total; e += PV_GROUPS) {
            const PsfVolRec &r = list[e];
            if (!valid || r.lo[0] > i0 || r.hi[0] < i0 || r.lo[1] > i1 || r.hi[1] < i1 || r.lo[2] > i2 ||
                r.hi[2] < i2)
                continue;
            psf_kahan(s, c, RULE == GT_PSF_SAMPLED ? psf_voxel_sampled<D>(r, yc) : psf_voxel<D>(r, yc, hv));
        }
        __syncthreads();
    }
    // groups' sums (value minus compensation), added in group order
    float *red = reinterpret_cast<float *>(list);
    static_assert(sizeof(list) >= PV_THREADS * sizeof(float), "reduction buffer");
    red[threadIdx.x] = s - c;
    __syncthreads();
    if (grp == 0 && valid) {
        float ts = 0.0f, tc = 0.0f;
        for (int g = 0; g < PV_GROUPS; g++) psf_kahan(ts, tc, red[g * 256 + t]);
        u[((size_t)i0 * V.m1 + i1) * V.m2 + i2] = ts;
    }
}

__global__ void psf_fill_kernel(float *__restrict__ p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

}  // namespace gt

using namespace gt;

static size_t psf_rec_bytes(const gt_atoms *a, const gt_geom *g) { return align_up((size_t)a->n * g->nA * sizeof(PsfRec)); }

extern "C" size_t gt_psf_forward_workspace(const gt_atoms *atoms, const gt_geom *geom)
{
    if (!atoms || !geom) return 0;
    return psf_rec_bytes(atoms, geom);
}

static int psf_prepare(const gt_atoms *atoms, const gt_geom *geom, int rule, void *ws, size_t ws_bytes, size_t need,
                       hipStream_t st, const char *who)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_geom(geom, atoms->dim))) return rc;
    GT_CHECK_ARG(rule == GT_PSF_EXACT || rule == GT_PSF_SAMPLED,
                 "%s: rule must be %d (pixel integral) or %d (the reference's samples), got %d", who, GT_PSF_EXACT,
                 GT_PSF_SAMPLED, rule);
    GT_CHECK_ARG(geom->sorted == 1, "%s: the detector coordinates must ascend (uniform pixels)", who);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("%s: workspace of %zu bytes given, %zu needed", who, ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "%s: workspace must be 64-byte aligned", who);
    GT_CHECK_ARG(geom->n0 >= 2 && (atoms->dim == 2 || geom->n1 >= 2),
                 "%s: pixel integration needs at least two detector coordinates per axis", who);
    if (atoms->n == 0) return 0;
    PsfRec *recs = (PsfRec *)ws;
    const long pairs = (long)atoms->n * geom->nA;
    const dim3 grid((unsigned)((pairs + 255) / 256));
    if (atoms->dim == 3)
        hipLaunchKernelGGL(psf_setup_kernel<3>, grid, dim3(256), 0, st, *atoms, *geom, recs, rule);
    else
        hipLaunchKernelGGL(psf_setup_kernel<2>, grid, dim3(256), 0, st, *atoms, *geom, recs, rule);
    GT_HIP(hipGetLastError());
    return 0;
}

extern "C" int gt_psf_forward_rule(const gt_atoms *atoms, const gt_geom *geom, int rule, float *R, void *ws,
                                   size_t ws_bytes, void *stream)
{
    int rc;
    hipStream_t st = (hipStream_t)stream;
    if (!atoms || !geom) {
        set_error("gt_psf_forward: atoms or geometry is NULL");
        return GT_EINVAL;
    }
    if ((rc = psf_prepare(atoms, geom, rule, ws, ws_bytes, gt_psf_forward_workspace(atoms, geom), st,
                          "gt_psf_forward")))
        return rc;
    GT_CHECK_ARG(R != nullptr, "gt_psf_forward: R is NULL");
    const int n1 = atoms->dim == 3 ? geom->n1 : 1;
    if (atoms->n == 0) {
        const size_t n = (size_t)geom->nA * geom->n0 * n1;
        hipLaunchKernelGGL(psf_fill_kernel, dim3((unsigned)std::min<size_t>((n + 255) / 256, 65535)), dim3(256), 0, st,
                           R, n);
        GT_HIP(hipGetLastError());
        return 0;
    }
    const PsfRec *recs = (const PsfRec *)ws;
    prof_begin(st, GT_K_FORWARD);
    if (atoms->dim == 3) {
        const int e0 = ceil_log2((geom->n0 + 15) / 16), e1 = ceil_log2((n1 + 15) / 16);
        GT_CHECK_ARG(((unsigned long long)geom->nA << (e0 + e1)) < (1ull << 31), "gt_psf_forward: too many tiles");
        const dim3 grid((unsigned)((unsigned long long)geom->nA << (e0 + e1)));
        if (rule == GT_PSF_SAMPLED)
            hipLaunchKernelGGL((psf_forward_kernel<2, GT_PSF_SAMPLED>), grid, dim3(256), 0, st, recs, atoms->n,
                               geom->nA, geom->n0, n1, geom->p0, geom->p1, R, e0, e1);
        else
            hipLaunchKernelGGL((psf_forward_kernel<2, GT_PSF_EXACT>), grid, dim3(256), 0, st, recs, atoms->n, geom->nA,
                               geom->n0, n1, geom->p0, geom->p1, R, e0, e1);
    } else {
        const dim3 grid((unsigned)((geom->n0 + 255) / 256), (unsigned)geom->nA);
        if (rule == GT_PSF_SAMPLED)
            hipLaunchKernelGGL((psf_forward_kernel<1, GT_PSF_SAMPLED>), grid, dim3(256), 0, st, recs, atoms->n,
                               geom->nA, geom->n0, 1, geom->p0, geom->p0, R, 0, 0);
        else
            hipLaunchKernelGGL((psf_forward_kernel<1, GT_PSF_EXACT>), grid, dim3(256), 0, st, recs, atoms->n, geom->nA,
                               geom->n0, 1, geom->p0, geom->p0, R, 0, 0);
    }
    prof_end(st, GT_K_FORWARD);
    GT_HIP(hipGetLastError());
    return 0;
}

extern "C" int gt_psf_forward(const gt_atoms *atoms, const gt_geom *geom, float *R, void *ws, size_t ws_bytes,
                              void *stream)
{
    return gt_psf_forward_rule(atoms, geom, GT_PSF_EXACT, R, ws, ws_bytes, stream);
}

extern "C" size_t gt_psf_derivs_workspace(const gt_atoms *atoms, const gt_geom *geom)
{
    if (!atoms || !geom) return 0;
    return psf_rec_bytes(atoms, geom) + align_up((size_t)atoms->n * geom->nA * PSF_MS * sizeof(float)) +
           align_up((size_t)geom->nA * sizeof(double));
}

template <int K, int MODE>
static void psf_launch_moments(int rule, dim3 mg, hipStream_t st, const PsfRec *recs, int N, int nA, int n0, int n1,
                               const float *p0, const float *p1, const float *C, int order, float *mom)
{
    if (rule == GT_PSF_SAMPLED)
        hipLaunchKernelGGL((psf_moments_kernel<K, MODE, GT_PSF_SAMPLED>), mg, dim3(256), 0, st, recs, N, nA, n0, n1, p0,
                           p1, C, order, mom);
    else
        hipLaunchKernelGGL((psf_moments_kernel<K, MODE, GT_PSF_EXACT>), mg, dim3(256), 0, st, recs, N, nA, n0, n1, p0,
                           p1, C, order, mom);
}

extern "C" int gt_psf_derivs_rule(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order,
                                  int rule, double *f, double *g, double *H, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    hipStream_t st = (hipStream_t)stream;
    if (!atoms || !geom) {
        set_error("gt_psf_derivs: atoms or geometry is NULL");
        return GT_EINVAL;
    }
    GT_CHECK_ARG(mode == GT_MODE_LINEAR || mode == GT_MODE_L2, "gt_psf_derivs: mode must be 0 (linear) or 1 (L2), got %d",
                 mode);
    GT_CHECK_ARG(order == 1 || order == 2, "gt_psf_derivs: order must be 1 or 2, got %d", order);
    GT_CHECK_ARG(C != nullptr || atoms->n == 0, "gt_psf_derivs: C is NULL");
    GT_CHECK_ARG(g != nullptr || atoms->n == 0, "gt_psf_derivs: g is NULL");
    if ((rc = psf_prepare(atoms, geom, rule, ws, ws_bytes, gt_psf_derivs_workspace(atoms, geom), st,
                          "gt_psf_derivs")))
        return rc;
    if (atoms->n == 0) return 0;
    char *w = (char *)ws;
    const PsfRec *recs = (const PsfRec *)w;
    float *mom = (float *)(w + psf_rec_bytes(atoms, geom));
    double *sumC2 = (double *)(w + psf_rec_bytes(atoms, geom) +
                               align_up((size_t)atoms->n * geom->nA * PSF_MS * sizeof(float)));
    const int n1 = atoms->dim == 3 ? geom->n1 : 1;
    const long pairs = (long)atoms->n * geom->nA;
    const dim3 mg((unsigned)((pairs + 3) / 4));
    if (mode == GT_MODE_L2)
        hipLaunchKernelGGL(psf_sumsq_kernel, dim3(geom->nA), dim3(256), 0, st, C, (long)geom->n0 * n1, sumC2);
    prof_begin(st, GT_K_MOMENTS);
    if (atoms->dim == 3) {
        if (mode == GT_MODE_L2)
            psf_launch_moments<2, GT_MODE_L2>(rule, mg, st, recs, atoms->n, geom->nA, geom->n0, n1, geom->p0, geom->p1,
                                              C, order, mom);
        else
            psf_launch_moments<2, GT_MODE_LINEAR>(rule, mg, st, recs, atoms->n, geom->nA, geom->n0, n1, geom->p0,
                                                  geom->p1, C, order, mom);
    } else {
        if (mode == GT_MODE_L2)
            psf_launch_moments<1, GT_MODE_L2>(rule, mg, st, recs, atoms->n, geom->nA, geom->n0, 1, geom->p0, geom->p0,
                                              C, order, mom);
        else
            psf_launch_moments<1, GT_MODE_LINEAR>(rule, mg, st, recs, atoms->n, geom->nA, geom->n0, 1, geom->p0,
                                                  geom->p0, C, order, mom);
    }
    prof_end(st, GT_K_MOMENTS);
    prof_begin(st, GT_K_FINISH);
    const int wh = order > 1;
    if (atoms->dim == 3) {
        if (mode == GT_MODE_L2)
            hipLaunchKernelGGL((psf_finish_kernel<3, GT_MODE_L2>), dim3(atoms->n), dim3(64), 0, st, *atoms, *geom,
                               recs, mom, sumC2, wh, f, g, H);
        else
            hipLaunchKernelGGL((psf_finish_kernel<3, GT_MODE_LINEAR>), dim3(atoms->n), dim3(64), 0, st, *atoms,
                               *geom, recs, mom, sumC2, wh, f, g, H);
    } else {
        if (mode == GT_MODE_L2)
            hipLaunchKernelGGL((psf_finish_kernel<2, GT_MODE_L2>), dim3(atoms->n), dim3(64), 0, st, *atoms, *geom,
                               recs, mom, sumC2, wh, f, g, H);
        else
            hipLaunchKernelGGL((psf_finish_kernel<2, GT_MODE_LINEAR>), dim3(atoms->n), dim3(64), 0, st, *atoms,
                               *geom, recs, mom, sumC2, wh, f, g, H);
    }
    prof_end(st, GT_K_FINISH);
    GT_HIP(hipGetLastError());
    return 0;
}

extern "C" int gt_psf_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order,
                             double *f, double *g, double *H, void *ws, size_t ws_bytes, void *stream)
{
    return gt_psf_derivs_rule(atoms, geom, C, mode, order, GT_PSF_EXACT, f, g, H, ws, ws_bytes, stream);
}

extern "C" size_t gt_psf_volume_workspace(const gt_atoms *atoms, const gt_vol *vol)
{
    (void)vol;
    if (!atoms) return 0;
    return align_up((size_t)atoms->n * sizeof(PsfVolRec));
}

extern "C" int gt_psf_volume_rule(const gt_atoms *atoms, const gt_vol *vol, int rule, float *u, void *ws,
                                  size_t ws_bytes, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms))) return rc;
    GT_CHECK_ARG(rule == GT_PSF_EXACT || rule == GT_PSF_SAMPLED,
                 "gt_psf_volume: rule must be %d (voxel integral) or %d (the reference's samples), got %d",
                 GT_PSF_EXACT, GT_PSF_SAMPLED, rule);
    GT_CHECK_ARG(vol != nullptr && vol->dim == atoms->dim, "gt_psf_volume: volume dimension must match the atoms");
    GT_CHECK_ARG(vol->m0 >= 1 && vol->m1 >= 1 && vol->m2 >= 1, "gt_psf_volume: grid sizes must be >= 1");
    GT_CHECK_ARG(vol->dim == 3 || vol->m2 == 1, "gt_psf_volume: 2D volumes have m2 == 1");
    GT_CHECK_ARG(vol->sorted == 1, "gt_psf_volume: the coordinate vectors must ascend (uniform voxels)");
    GT_CHECK_ARG(vol->y0 && vol->y1 && (vol->dim == 2 || vol->y2), "gt_psf_volume: coordinate vectors missing");
    GT_CHECK_ARG(u != nullptr, "gt_psf_volume: u is NULL");
    const size_t need = gt_psf_volume_workspace(atoms, vol);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_psf_volume: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "gt_psf_volume: workspace must be 64-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    GT_CHECK_ARG(vol->m0 >= 2 && vol->m1 >= 2 && (vol->dim == 2 || vol->m2 >= 2),
                 "gt_psf_volume: voxel integration needs at least two coordinates per axis");
    const size_t nout = (size_t)vol->m0 * vol->m1 * vol->m2;
    if (atoms->n == 0) {
        hipLaunchKernelGGL(psf_fill_kernel, dim3((unsigned)std::min<size_t>((nout + 255) / 256, 65535)), dim3(256), 0,
                           st, u, nout);
        GT_HIP(hipGetLastError());
        return 0;
    }
    PsfVolRec *recs = (PsfVolRec *)ws;
    const dim3 sg((unsigned)((atoms->n + 255) / 256));
    // bricks numbered by Morton index over 2^e0 x 2^e1 x 2^e2 centre-out ranks
    int nb[3];
    if (vol->dim == 3) {
        hipLaunchKernelGGL(psf_vol_setup_kernel<3>, sg, dim3(256), 0, st, *atoms, *vol, recs, rule);
        nb[0] = (vol->m0 + 3) / 4, nb[1] = (vol->m1 + 7) / 8, nb[2] = (vol->m2 + 7) / 8;
    } else {
        hipLaunchKernelGGL(psf_vol_setup_kernel<2>, sg, dim3(256), 0, st, *atoms, *vol, recs, rule);
        nb[0] = (vol->m0 + 15) / 16, nb[1] = (vol->m1 + 15) / 16, nb[2] = 1;
    }
    const int e[3] = {ceil_log2(nb[0]), ceil_log2(nb[1]), ceil_log2(nb[2])};
    GT_CHECK_ARG(e[0] + e[1] + e[2] <= 30, "gt_psf_volume: grid too large");
    const unsigned blocks = 1u << (e[0] + e[1] + e[2]);
    prof_begin(st, GT_K_VOLUME);
    if (vol->dim == 3) {
        if (rule == GT_PSF_SAMPLED)
            hipLaunchKernelGGL((psf_volume_kernel<3, GT_PSF_SAMPLED>), dim3(blocks), dim3(PV_THREADS), 0, st, recs,
                               atoms->n, *vol, u, e[0], e[1], e[2]);
        else
            hipLaunchKernelGGL((psf_volume_kernel<3, GT_PSF_EXACT>), dim3(blocks), dim3(PV_THREADS), 0, st, recs,
                               atoms->n, *vol, u, e[0], e[1], e[2]);
    } else {
        if (rule == GT_PSF_SAMPLED)
            hipLaunchKernelGGL((psf_volume_kernel<2, GT_PSF_SAMPLED>), dim3(blocks), dim3(PV_THREADS), 0, st, recs,
                               atoms->n, *vol, u, e[0], e[1], e[2]);
        else
            hipLaunchKernelGGL((psf_volume_kernel<2, GT_PSF_EXACT>), dim3(blocks), dim3(PV_THREADS), 0, st, recs,
                               atoms->n, *vol, u, e[0], e[1], e[2]);
    }
    prof_end(st, GT_K_VOLUME);
    GT_HIP(hipGetLastError());
    return 0;
}

extern "C" int gt_psf_volume(const gt_atoms *atoms, const gt_vol *vol, float *u, void *ws, size_t ws_bytes,
                             void *stream)
{
    return gt_psf_volume_rule(atoms, vol, GT_PSF_EXACT, u, ws, ws_bytes, stream);
}
