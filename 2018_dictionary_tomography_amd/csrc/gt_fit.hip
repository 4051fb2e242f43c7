// gt_fit.hip -- device-resident damped-Newton fitter.
//
// Restates NewtonGaussian.linesearch (NewtonGaussian.py:18-109) with
// GD_lib._step (GD_lib.py:492-536), and NewtonGaussian.linesearch_block
// (NewtonGaussian.py:112-237), for the l2 fidelity and any per-atom
// regulariser of gt_reg.h.  For each atom j of the outer iteration's
// permutation:
//
//   1. C = data + R_j - R is never formed: the L2 derivatives of atom j
//      against C equal the derivatives of 1/2||R - data||^2 w.r.t. atom j,
//      so the pixel moments are taken against the stored residual R - data.
//   2. (g, H) of atom j: per-angle FP32 pixel moments + FP64 contraction,
//      plus the regulariser's gradient and Hessian (NewtonGaussian.py:61-66).
//   3. dx = solve(H + I/h_j, -g) in FP64 with partial pivoting (numpy.linalg.
//      solve, NewtonGaussian.py:73-76); singular -> the fit stops (done = 2).
//      Block variant: the I, x and r blocks in turn, each solved on its own
//      diagonal block of H + I/h_{t,j}; a block whose system is singular (or
//      whose I pivot is <= 1e-8) is skipped (NewtonGaussian.py:154-176).
//   4. Projections: _step's (I >= 0, rows of r with a negative diagonal
//      negated); the block variant also clamps x to [-0.99, 0.99] (:167).
//   5. Energy change: the change of the committed FP32 residual's
//      1/2||.||^2, accumulated in FP64 over the union of the old and new
//      footprints, plus reg(new_j) - reg(old_j); reject iff the new total
//      exceeds the old (the reference's E[jj] > E[jj-1] test), and reject a
//      proposal whose fidelity change is not finite.
//   6. Accepted: atom written back, residual updated over the same pixels.
//      Step size h x2 (block: x1.3) / /10 clamped to [1e-4, 1e4], convergence
//      flags as in NewtonGaussian.py:78-98 and :205-221.
//
// A slot is advanced by a group of G workgroups of 16 waves (G = 1 for
// gt_fit_run).  Member m owns angles m, m + G, ...: their detector Gaussians
// (LDS), pixel moments and residual slices.  Every
// cross-angle quantity -- the 10 + 55 gradient/Hessian entries and the energy
// change -- is first formed per angle in the slot's scratch and then summed
// over angles in one fixed order by every member, so all members take the
// same decisions from the same numbers and the result does not depend on G.
// The members meet at two device-wide barriers per atom update (after the
// per-angle contractions and after the per-angle energy changes); a barrier
// wait is bounded in time.  The solve and the accept logic run redundantly in
// every member; member 0 writes the progress fields back at the end.
//
// Pixel phases give one (atom, angle) pair to each wave.  A wave contracts
// the moments of its angle as soon as it has reduced them (the FP64 work of
// one pair spread over the wave's lanes, see prep1..prep3), so the latency of
// that FP64 chain overlaps the pixel work of the other waves.
#include <mutex>

#include "gt_pixel.h"
#include "gt_reg.h"

namespace gt {

constexpr int FIT_THREADS = 1024;
constexpr int FIT_WAVES = FIT_THREADS / 64;
constexpr int NAMAX = GT_FIT_MAX_ANGLES;
static_assert(NAMAX % 64 == 0, "per-angle sums take whole waves of angles");
constexpr int DETMAX = GT_FIT_MAX_DETECTOR;
constexpr int NVS = 65;                          // stride of a per-angle gradient/Hessian record
constexpr int FIT_UNR = 8;                       // detector rows in flight per lane in the pixel passes
constexpr int FIT_TROWS = 192;                   // rows of a wave's row table (moments, energy and commit passes)
constexpr int FIT_BOXW = 64;                     // columns per chunk of the moments walk
constexpr double FIT_EPS = 1e-4;                 // NewtonGaussian.py:21
constexpr double GD_EPS = 1e-8;                  // GD_lib.py:96
constexpr long long FIT_WAIT_TICKS = 200000000;  // barrier wait limit: 2 s of the 100 MHz clock

// Slot scratch layout (bytes): per-angle records double[nA][NVS] | per-angle
// energy changes double[nA] | barrier words.
__host__ __device__ inline size_t fit_off_dE(int nA) { return align_up((size_t)nA * NVS * sizeof(double)); }
__host__ __device__ inline size_t fit_off_bar(int nA) { return fit_off_dE(nA) + align_up((size_t)nA * sizeof(double)); }
__host__ __device__ inline size_t fit_scratch_size(int nA) { return fit_off_bar(nA) + ALIGN; }

// Progress of a fit, held in LDS by every member during a launch.
struct FitCtl {
    int it, pos, inner, jj, brk0, brk1, done, blk, ninf, pad;
    double E, F, Rsum;
    long long updates, accepted;
};

// Per-wave LDS of the fused moments + contraction pass: one pair at a time.
template <int D> struct WaveC {
    PairFin<D> F;
    PairAlg<D> a;
    PairGeom<D> g;
    double S1[15], S2[15];
};

template <int D> struct FitLDS {
    PairRec rec_old[NAMAX];  // indexed by the member's own angle number
    PairRec rec_new[NAMAX];
    WaveC<D> wc[FIT_WAVES];
    float p0[DETMAX], p1[DETMAX];
    float geo[NAMAX][9];               // own angles' T, w0, w1 (as load_angle reads them), once per launch
    float4 tab[FIT_WAVES][FIT_TROWS];  // per-wave row tables of the pixel passes
    double th_old[10], th_new[10];
    double g[10], H[100];
    double A[100], b[10], x[10];
    double A2[36], b2[6];     // diagonal block of the block variant
    double rg[10], rH[100];   // regulariser gradient / Hessian of the current atom
    double rv_old, rv_new;    // regulariser terms of the current atom and of the proposal
    double dxnorm, hj;
    FitCtl c;
    unsigned gen;
    int accept, j, stop, abort_, skip, moved;
    int next[3];  // next own-angle number of the moments, energy and commit passes
};

// The wave's next own angle of a pixel pass: waves take angles as they
// finish the previous one (the angles' boxes differ in size), so a pass ends
// when the work runs out rather than when the wave with the largest boxes
// does.  Every angle's result goes to a slot of its own and the sums over
// angles run in angle order, so the results do not depend on which wave
// took which angle.
__device__ inline int next_angle(int *ctr, int lane)
{
    int k = 0;
    if (lane == 0) k = atomicAdd(ctr, 1);
    return __builtin_amdgcn_readfirstlane(k);
}

// The geometry of angle a (T, w0 and, for a 3D atom, w1) as load_angle reads
// it, into 9 floats of LDS; and back into a PairGeom, with the same values.
template <int D> __device__ inline void stash_angle(const gt_geom &G, int a, float *o)
{
    const int L = G.lenT;
    for (int i = 0; i < 9; i++) o[i] = 0.0f;
    for (int i = 0; i < L; i++) {
        o[i] = G.T[(size_t)a * L + i];
        o[3 + i] = G.w0[(size_t)a * L + i];
    }
    if (D == 3 && L == 3)
        for (int i = 0; i < 3; i++) o[6 + i] = G.w1[(size_t)a * 3 + i];
}
template <int D> __device__ inline void angle_from_stash(const float *o, int L, PairGeom<D> &g)
{
#pragma unroll
    for (int i = 0; i < 3; i++) {
        g.T[i] = o[i];
        g.G[0][i] = o[3 + i];
        g.G[1][i] = 0.0;
    }
    if (D == 3) {
        if (L == 3) {
#pragma unroll
            for (int i = 0; i < 3; i++) g.G[1][i] = o[6 + i];
        } else {
            g.G[1][2] = 1.0;  // 2.5D, as load_angle
        }
    }
}

// theta (FP64 values of the FP32 atom) -> PairGeom atom part
template <int D> __device__ inline void atom_from_theta(const double *th, int iso, PairGeom<D> &g)
{
    constexpr int NR = Dims<D>::NR;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) g.U[i][k] = 0.0;
    g.I = th[0];
    for (int i = 0; i < 3; i++) g.x[i] = (i < D) ? th[1 + i] : 0.0;
    if (iso) {
        for (int i = 0; i < D; i++) g.U[i][i] = th[1 + D];
    } else {
        for (int k = 0; k < NR; k++) g.U[slot_row(D, k)][slot_col(D, k)] = th[1 + D + k];
    }
}

__device__ inline gt_geom slot_geom(const gt_fit_slot &S, int dim)
{
    gt_geom G;
    G.dim = dim;
    G.lenT = S.lenT;
    G.nA = S.nA;
    G.n0 = S.n0;
    G.n1 = S.n1;
    G.sorted = S.sorted;
    G.T = S.T;
    G.w0 = S.w0;
    G.w1 = S.w1;
    G.p0 = S.p0;
    G.p1 = S.p1;
    return G;
}

// Union of two (possibly empty) boxes.
__device__ inline void box_union(const PairRec &a, const PairRec &b, int K, int *lo, int *hi)
{
    const bool ea = a.lo[0] > a.hi[0] || (K == 2 && a.lo[1] > a.hi[1]);
    const bool eb = b.lo[0] > b.hi[0] || (K == 2 && b.lo[1] > b.hi[1]);
    for (int k = 0; k < 2; k++) {
        if (ea && eb) {
            lo[k] = 1;
            hi[k] = 0;
        } else if (ea) {
            lo[k] = b.lo[k];
            hi[k] = b.hi[k];
        } else if (eb) {
            lo[k] = a.lo[k];
            hi[k] = a.hi[k];
        } else {
            lo[k] = min(a.lo[k], b.lo[k]);
            hi[k] = max(a.hi[k], b.hi[k]);
        }
    }
}

__device__ inline float pair_value(const PairRec &rec, float p0v, float p1v)
{
    const float dp0 = (p0v - rec.c_hi[0]) - rec.c_lo[0];
    const float dp1 = (p1v - rec.c_hi[1]) - rec.c_lo[1];
    const float q = pair_q(rec.q, dp0, dp1);
    return (q >= (float)KAPPA) ? rec.alpha * __builtin_amdgcn_exp2f(q) : 0.0f;
}

// Replacing rec_old by rec_new over the union box, both values evaluated per
// pixel with the same FP32 expressions in every mode; returns the lane's FP64
// energy-change partial (modes 0 and 2): the sum of (new - old)(new + old),
// each term exact in FP64, halved once at the end (the same bits as halving
// every term: scaling by 2 commutes with rounding).
//   mode 0  energy change only (residual untouched)
//   mode 1  commit: residual += new - old
//   mode 2  speculative commit: energy change, residual updated in place, and
//           the old residual of every union-box pixel saved in Sa (same layout)
//           so that pair_restore can undo a rejected proposal exactly
// On a 2D detector the union box is taken in windows of at most FIT_TROWS
// rows; per window the wave first writes the row table dt[r] = (dp0 of the
// old record, dp0 of the new one) of its rows (the expressions a pixel would
// evaluate), so a pixel reads its row terms instead of forming them.
template <int K, int mode>
__device__ inline double pair_delta4(const PairRec &ro, const PairRec &rn, const float *p0, const float *p1,
                                     gmem<float> *__restrict__ Ra, gmem<float> *__restrict__ Sa, int n1, int lane,
                                     float2 *dt)
{
    int lo[2], hi[2];
    box_union(ro, rn, K, lo, hi);
    const int W = (K == 2) ? hi[1] - lo[1] + 1 : 1;
    const int H = hi[0] - lo[0] + 1;
    const long total = (W > 0 && H > 0) ? (long)W * H : 0;
    double acc = 0.0;
    if (total == 0) return acc;
    if (K == 2) {
        // column-owning walk (ColChunk): per-column terms of both Gaussians once
        // per lane; 32-bit element offsets within the angle's slab
        const float kap = (float)KAPPA;
        for (int w0 = 0; w0 < H; w0 += FIT_TROWS) {
            const int Hw = min(FIT_TROWS, H - w0);
            if (w0 > 0) wave_sync();  // the previous window's table is no longer read
            for (int r = lane; r < Hw; r += 64) {
                const float pv0 = p0[lo[0] + w0 + r];
                dt[r] = make_float2((pv0 - ro.c_hi[0]) - ro.c_lo[0], (pv0 - rn.c_hi[0]) - rn.c_lo[0]);
            }
            wave_sync();
            for (int cb = 0; cb < W; cb += 64) {
                ColChunk cc;
                cc.init(lane, W, cb);
                if (!cc.act) continue;
                const int i1 = lo[1] + cb + cc.c;
                const float pv1 = p1[i1];
                const float d1o = (pv1 - ro.c_hi[1]) - ro.c_lo[1], d1n = (pv1 - rn.c_hi[1]) - rn.c_lo[1];
                const float rlo = fmaf(ro.q[1], d1o, ro.q[3]), rln = fmaf(rn.q[1], d1n, rn.q[3]);
                const float clo = fmaf(fmaf(ro.q[2], d1o, ro.q[4]), d1o, ro.q[5]);
                const float cln = fmaf(fmaf(rn.q[2], d1n, rn.q[4]), d1n, rn.q[5]);
                const unsigned off0 = (unsigned)(lo[0] + w0) * (unsigned)n1 + (unsigned)i1;
                for (int r0 = cc.ro; r0 < Hw; r0 += FIT_UNR * cc.R) {
                    float old[FIT_UNR];
                    float2 d0[FIT_UNR];
                    bool ok[FIT_UNR];
                    unsigned off[FIT_UNR];
#pragma unroll
                    for (int u = 0; u < FIT_UNR; u++) {
                        const int r = r0 + u * cc.R;
                        ok[u] = r < Hw;
                        const unsigned rr = ok[u] ? (unsigned)r : 0u;
                        off[u] = off0 + rr * (unsigned)n1;
                        old[u] = Ra[off[u]];
                        d0[u] = dt[rr];
                    }
#pragma unroll
                    for (int u = 0; u < FIT_UNR; u++) {
                        const float d0o = d0[u].x, d0n = d0[u].y;
                        const float qo = fmaf(fmaf(ro.q[0], d0o, rlo), d0o, clo);
                        const float qn = fmaf(fmaf(rn.q[0], d0n, rln), d0n, cln);
                        const float vo = (qo >= kap) ? ro.alpha * __builtin_amdgcn_exp2f(qo) : 0.0f;
                        const float vn = (qn >= kap) ? rn.alpha * __builtin_amdgcn_exp2f(qn) : 0.0f;
                        const float dR = vn - vo;
                        if (mode == 2 && ok[u]) Sa[off[u]] = old[u];
                        if (ok[u] && dR != 0.0f) {
                            const float nw = old[u] + dR;
                            if (mode != 0) Ra[off[u]] = nw;
                            if (mode != 1) acc += ((double)nw - (double)old[u]) * ((double)nw + (double)old[u]);
                        }
                    }
                }
            }
        }
        wave_sync();  // the table is rewritten for the wave's next pair
        return 0.5 * acc;
    }
    Track4 tr;
    tr.init(lane, W);
    for (long k = lane; k < total; k += 256) {
        float old[4], pv0[4], pv1[4];
        size_t idx[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            ok[u] = k + 64 * u < total;
            const int i0 = lo[0] + (ok[u] ? tr.r[u] : 0);
            const int i1 = (K == 2) ? lo[1] + (ok[u] ? tr.c[u] : 0) : 0;
            idx[u] = (size_t)i0 * n1 + i1;
            old[u] = Ra[idx[u]];
            pv0[u] = p0[i0];
            pv1[u] = (K == 2) ? p1[i1] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float dR = pair_value(rn, pv0[u], pv1[u]) - pair_value(ro, pv0[u], pv1[u]);
            if (mode == 2 && ok[u]) Sa[idx[u]] = old[u];
            if (ok[u] && dR != 0.0f) {
                const float nw = old[u] + dR;
                if (mode != 0) Ra[idx[u]] = nw;
                if (mode != 1) acc += ((double)nw - (double)old[u]) * ((double)nw + (double)old[u]);
            }
        }
        tr.step();
    }
    return 0.5 * acc;
}

// Undo a speculative commit: the union box of the residual back from Sa.
template <int K>
__device__ inline void pair_restore(const PairRec &ro, const PairRec &rn, gmem<float> *__restrict__ Ra,
                                    const gmem<float> *__restrict__ Sa, int n1, int lane)
{
    int lo[2], hi[2];
    box_union(ro, rn, K, lo, hi);
    const int W = (K == 2) ? hi[1] - lo[1] + 1 : 1;
    const int H = hi[0] - lo[0] + 1;
    if (W <= 0 || H <= 0) return;
    for (int k = lane; k < W * H; k += 64) {
        const int r = k / W, c = k - (k / W) * W;
        const unsigned off = (unsigned)(lo[0] + r) * (unsigned)n1 + (unsigned)(K == 2 ? lo[1] + c : 0);
        Ra[off] = Sa[off];
    }
}

// pair_alg with detector direction b on lane b (b = 0..K; other lanes idle),
// the same expressions as pair_alg.  B is not needed by the contraction.
template <int D>
__device__ inline void wave_pair_alg(const PairGeom<D> &g, const double *pstar, PairAlg<D> &a, int lane)
{
    constexpr int K = Dims<D>::K;
    if (lane > K) return;
    double t[3] = {0.0, 0.0, 0.0}, s = 0.0;
    matvec_upper<D>(g.U, g.T, t);
    for (int i = 0; i < D; i++) s += t[i] * t[i];
    const int b = lane;
    double h[3] = {0.0, 0.0, 0.0}, u[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < D; i++) {
        if (b < K) {
            h[i] = g.G[b][i];
        } else {
            double y = 0.0;
            for (int k = 0; k < K; k++) y += pstar[k] * g.G[k][i];
            h[i] = y - g.x[i];
        }
    }
    matvec_upper<D>(g.U, h, u);
    double ip = 0.0;
    for (int i = 0; i < D; i++) ip += u[i] * t[i];
    const double mu = ip / s;
    for (int i = 0; i < 3; i++) {
        a.h[b][i] = h[i];
        a.u[b][i] = u[i];
        a.ut[b][i] = (i < D) ? u[i] - mu * t[i] : 0.0;
        a.ht[b][i] = (i < D) ? h[i] - mu * g.T[i] : 0.0;
    }
    a.mu[b] = mu;
    if (lane == 0) {
        for (int i = 0; i < 3; i++) a.t[i] = t[i];
        a.s = s;
        a.alpha = g.I / sqrt(s);
    }
}

// Value of v in lane l (l the same in every lane): two v_readlane, no LDS.
__device__ inline double readlane_f64(double v, int l)
{
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)bits, l);
    const int hi = __builtin_amdgcn_readlane((int)(bits >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// Solve A x = b (n x n, row-major in LDS) with partial pivoting on one wave,
// lane i holding row i; rows are broadcast with lane reads.  Pivot = first row
// of largest magnitude (LAPACK i*amax), rows swapped as in LAPACK getrf.
// Returns false if a pivot is zero (or not a number).
template <int n> __device__ inline bool wave_solve(const double *A, const double *b, double *x, int lane)
{
    const bool own = lane < n;
    double a[n], bb = own ? b[lane] : 0.0;
#pragma unroll
    for (int c = 0; c < n; c++) a[c] = own ? A[lane * n + c] : 0.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < n; k++) {
        const double colk = fabs(a[k]);
        double best = readlane_f64(colk, k);
        int p = k;
#pragma unroll
        for (int r = k + 1; r < n; r++) {
            const double v = readlane_f64(colk, r);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (!(best > 0.0)) ok = false;
        if (p != k) {
#pragma unroll
            for (int c = 0; c < n; c++) {
                const double ak = readlane_f64(a[c], k), ap = readlane_f64(a[c], p);
                if (lane == k) a[c] = ap;
                if (lane == p) a[c] = ak;
            }
            const double bk = readlane_f64(bb, k), bp = readlane_f64(bb, p);
            if (lane == k) bb = bp;
            if (lane == p) bb = bk;
        }
        double pk[n];
#pragma unroll
        for (int c = 0; c < n; c++) pk[c] = readlane_f64(a[c], k);
        const double pb = readlane_f64(bb, k);
        if (own && lane > k) {
            const double f = a[k] / pk[k];
#pragma unroll
            for (int c = 0; c < n; c++)
                if (c >= k) a[c] -= f * pk[c];
            bb -= f * pb;
        }
    }
    double xs[n];
#pragma unroll
    for (int i = n - 1; i >= 0; i--) {
        double s = bb;
#pragma unroll
        for (int c = 0; c < n; c++)
            if (c > i) s -= a[c] * xs[c];
        xs[i] = readlane_f64((lane == i) ? s / a[i] : 0.0, i);
    }
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < n; i++) x[i] = xs[i];
    return ok;
}

// wave_solve for the system sizes a fit of dimension D can ask for: all
// parameters (aniso P, iso 2 + D), or one block (x: D, r: NR or 1).
template <int D> __device__ inline bool solve_n(int n, const double *A, const double *b, double *x, int lane)
{
    switch (n) {
    case 1:
        return wave_solve<1>(A, b, x, lane);
    case D:
        return wave_solve<D>(A, b, x, lane);
    case Dims<D>::NR:
        return wave_solve<Dims<D>::NR>(A, b, x, lane);
    case 2 + D:
        return wave_solve<2 + D>(A, b, x, lane);
    case Dims<D>::P:
        return wave_solve<Dims<D>::P>(A, b, x, lane);
    default:
        return false;
    }
}

// Upper-triangle entry index e (0-based, row-major over i <= j) -> (i, j).
__device__ inline void upper_ij(int e, int P, int &i, int &j)
{
    i = 0;
    while (e >= P - i) {
        e -= P - i;
        i++;
    }
    j = i + e;
}

// Iso collapse: theta_iso index l -> aniso indices {l} (l <= D) or the D diagonal slots.
template <int D> __device__ inline double collapse_H(const double *H, int P, int i, int l)
{
    double s = 0.0;
    const int ni = (i <= D) ? 1 : D, nl = (l <= D) ? 1 : D;
    for (int u = 0; u < ni; u++)
        for (int v = 0; v < nl; v++) s += H[(i <= D ? i : 1 + D + u) * P + (l <= D ? l : 1 + D + v)];
    return s;
}

// Cross-workgroup traffic of a group (per-angle records, energy changes,
// barrier words) uses device-scope relaxed accesses: they are performed at the
// device's point of coherence (sc1), so no L2 write-back or invalidation is
// needed -- those would also evict the cached residuals of every other fit on
// the same XCD.  Ordering comes from waiting for completion (s_waitcnt) before
// a barrier is signalled, and from issuing the reads after it is observed.
__device__ inline void xput(double *p, double v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline double xget(const double *p)
{
    return __hip_atomic_load(const_cast<double *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned xgetu(unsigned *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void xputu(unsigned *p, unsigned v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void wait_all_memory() { __builtin_amdgcn_s_waitcnt(0); }

// Device-wide barrier of the G workgroups of one slot, called by thread 0 of
// each after every thread of its workgroup has completed its xput stores.
// bar[0] counts arrivals, bar[1] is the generation, bar[2] the abort flag.
// The wait is bounded: after FIT_WAIT_TICKS the slot is aborted.  Returns
// false if aborted.
__device__ inline bool group_barrier(unsigned *bar, int G, unsigned &gen)
{
    if (xgetu(bar + 2) != 0u) return false;
    const unsigned target = gen + 1u;
    const unsigned old = __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == (unsigned)G - 1u) {
        xputu(bar, 0u);
        wait_all_memory();  // the reset lands before anyone can pass and arrive again
        xputu(bar + 1, target);
    } else {
        const long long t0 = wall_clock64();
        while (xgetu(bar + 1) != target) {
            if (xgetu(bar + 2) != 0u) return false;
            if (wall_clock64() - t0 > FIT_WAIT_TICKS) {
                xputu(bar + 2, 1u);
                return false;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    }
    gen = target;
    return true;
}

// Advance the loop position after an update (or after a skipped block):
// inner iterations, the blocks of the block variant, atoms of the
// permutation, outer iterations and the convergence exit
// (NewtonGaussian.py:50-98, :131-222).
__device__ inline void fit_advance(FitCtl &c, const gt_fit_slot &S, bool skipped)
{
    if (S.block == 2) {
        // block GD: an inner iteration takes every block of the atom in turn
        // (GD_lib.py:133-170); BREAK[1] holds across its blocks
        if (++c.blk < S.gd_nblk) return;
        c.blk = 0;
        c.inner++;
        if (c.brk1 || c.inner >= S.max_inner) {
            c.inner = 0;
            c.pos++;
        }
    } else {
        bool next = skipped;
        if (!skipped) {
            c.inner++;
            next = c.brk1 || c.inner >= S.max_inner;
        }
        if (next) {
            c.inner = 0;
            if (S.block && ++c.blk < 3) {
                // next block of the same atom
            } else {
                c.blk = 0;
                c.pos++;
            }
        }
    }
    c.brk1 = 1;  // BREAK[1] = True at the start of every inner iteration
    if (c.pos >= S.N) {
        c.pos = 0;
        if (c.brk0 && c.it > S.min_iter) {
            c.done = 1;
        } else {
            c.it++;
            if (c.it >= S.max_outer) c.done = 1;
            c.brk0 = 1;
        }
    }
}

template <int D>
__global__ __launch_bounds__(FIT_THREADS) void fit_kernel(gt_fit_slot *__restrict__ slots, int nslots,
                                                         const gt_fit_block *__restrict__ blocks, int budget)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, P = Dm::P, NR = Dm::NR;
    constexpr int NV = P + P * (P + 1) / 2;
    static_assert(NV <= NVS, "per-angle record stride too small");
    __shared__ FitLDS<D> L;
    int sid = blockIdx.x, member = 0, G = 1, quota = 0;
    if (blocks) {
        const gt_fit_block e = blocks[blockIdx.x];
        sid = e.slot;
        member = e.member;
        G = e.group;
        quota = e.quota;
    }
    if (sid < 0 || sid >= nslots || G < 1 || member < 0 || member >= G) return;
    gt_fit_slot &S = slots[sid];
    if (!S.active || S.done) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int iso = S.iso;
    const int PI = iso ? 2 + D : P;  // parameters of the solve
    const int NRs = iso ? 1 : NR;
    const gt_geom Gm = slot_geom(S, D);
    const int N = S.N, nA = S.nA, n1 = S.n1;
    const size_t slab = (size_t)S.n0 * S.n1;
    const float *det[2] = {S.p0, S.p1};
    const int nd[2] = {S.n0, S.n1};
    char *scr = (char *)S.scratch;
    double *prec = (double *)scr;
    // A slot with a staging slab commits in the energy pass (no pass over the
    // residual follows an accepted update; a rejected one is undone from the
    // slab); without one, an accepted update is committed by a second pass.
    // The slab costs a store of every pixel of the union box; which is faster
    // depends on the load (the host decides per slot).  Both paths add the
    // same dR to the same old value, so results do not depend on it.
    gmem<float> *stage = (gmem<float> *)S.stage;
    double *dEa = (double *)(scr + fit_off_dE(nA));
    unsigned *bar = (unsigned *)(scr + fit_off_bar(nA));
    const int nOwn = member < nA ? (nA - member + G - 1) / G : 0;
    for (int i = tid; i < S.n0; i += FIT_THREADS) L.p0[i] = S.p0[i];
    if (K == 2)
        for (int i = tid; i < S.n1; i += FIT_THREADS) L.p1[i] = S.p1[i];
    for (int k = tid; k < nOwn; k += FIT_THREADS) stash_angle<D>(Gm, member + G * k, L.geo[k]);
    if (tid == 0) {
        FitCtl &c = L.c;
        c.it = S.it;
        c.pos = S.pos;
        c.inner = S.inner;
        c.jj = S.jj;
        c.brk0 = S.brk0;
        c.brk1 = S.brk1;
        c.done = S.done;
        c.E = S.E;
        c.updates = S.updates;
        c.accepted = S.accepted;
        c.blk = S.block ? S.blk : 0;
        c.ninf = S.ninf;
        c.F = S.F;
        c.Rsum = S.Rsum;
        L.gen = (G > 1) ? xgetu(bar + 1) : 0u;
        L.abort_ = 0;
    }
    __syncthreads();

    long long ticks[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long last = wall_clock64();
    const bool timing = member == 0 && S.phase_ticks != nullptr;
#define GT_FIT_TICK(k)                             \
    if (tid == 0 && timing) {                      \
        const long long now_ = wall_clock64();     \
        ticks[k] += now_ - last;                   \
        last = now_;                               \
    }
    // Publish this workgroup's global writes and meet the other members.
#define GT_FIT_MEET()                                                          \
    wait_all_memory();                                                         \
    __syncthreads();                                                           \
    if (G > 1) {                                                               \
        if (tid == 0 && !group_barrier(bar, G, L.gen)) L.abort_ = 1;           \
        __syncthreads();                                                       \
        if (L.abort_) break;                                                   \
    }

    const int my_budget = quota > 0 ? min(budget, quota) : (S.quota > 0 ? min(budget, S.quota) : budget);
    for (int step = 0; step < my_budget; step++) {
        // ---- choose the atom, load its parameters -----------------------
        if (tid == 0) {
            FitCtl &c = L.c;
            L.skip = 0;
            L.stop = c.done != 0 || c.it >= S.max_outer || c.pos >= N;
            if (!L.stop) {
                const int j = S.perm[(size_t)c.it * N + c.pos];
                if (j < 0 || j >= N) {
                    L.stop = 1;
                    c.done = 2;
                }
                L.j = j;
                L.next[0] = L.next[1] = L.next[2] = 0;
                if (!L.stop) {
                    L.th_old[0] = S.I[j];
                    for (int i = 0; i < D; i++) L.th_old[1 + i] = S.x[(size_t)j * D + i];
                    for (int k = 0; k < NRs; k++) L.th_old[1 + D + k] = S.r[(size_t)j * NRs + k];
                }
            }
        }
        __syncthreads();
        if (L.stop) break;
        const int j = L.j;
        // ---- phase A: detector Gaussians of the current atom (own angles) --
        for (int k = tid; k < nOwn; k += FIT_THREADS) {
            PairGeom<D> g;
            atom_from_theta<D>(L.th_old, iso, g);
            angle_from_stash<D>(L.geo[k], Gm.lenT, g);
            pair_setup<D>(g, det, nd, S.sorted, L.rec_old[k]);
        }
        __syncthreads();
        GT_FIT_TICK(0);
        // ---- phase B+C: per-angle pixel moments, then that angle's FP64
        //      gradient/Hessian record, both by the wave that took the angle --
        {
            WaveC<D> &W = L.wc[wave];
            for (int k = next_angle(&L.next[0], lane); k < nOwn; k = next_angle(&L.next[0], lane)) {
                const int a = member + G * k;
                const PairRec &rec = L.rec_old[k];
                float acc[MOMW];
#pragma unroll
                for (int v = 0; v < MOMW; v++) acc[v] = 0.0f;
                const gmem<float> *Ra = (gmem<float> *)S.resid + (size_t)a * slab;
                if (K == 2)
                    pair_moments_box<GT_MODE_RESID, FIT_UNR, FIT_TROWS, FIT_BOXW>(rec, L.p0, L.p1, Ra, n1, lane,
                                                                                  acc, L.tab[wave]);
                else
                    pair_moments4<K, GT_MODE_RESID>(rec, L.p0, L.p1, (const float *)Ra, n1, lane, acc);
                float mine = 0.0f;
#pragma unroll
                for (int v = 0; v < 30; v++) {
                    const float t = wave_sum_f32(acc[v]);
                    if (lane == v) mine = t;
                }
                const bool empty = rec.lo[0] > rec.hi[0] || (K == 2 && rec.lo[1] > rec.hi[1]);
                if (lane < 15)
                    W.S1[lane] = empty ? 0.0 : (double)mine;
                else if (lane < 30)
                    W.S2[lane - 15] = empty ? 0.0 : (double)mine;
                if (lane == 0) {
                    atom_from_theta<D>(L.th_old, iso, W.g);
                    angle_from_stash<D>(L.geo[k], Gm.lenT, W.g);
                }
                wave_sync();
                const double pstar[2] = {(double)rec.c_hi[0] + (double)rec.c_lo[0],
                                         (double)rec.c_hi[1] + (double)rec.c_lo[1]};
                wave_pair_alg<D>(W.g, pstar, W.a, lane);
                wave_sync();
                for (int i = lane; i < PrepCount<D>::N1; i += 64) prep1<D>(W.g, W.a, W.S1, W.S2, 1, i, W.F);
                wave_sync();
                if (lane < PrepCount<D>::N2) prep2<D>(W.g, W.a, lane, W.F);
                wave_sync();
                for (int i = lane; i < PrepCount<D>::N3 + PrepCount<D>::N4; i += 64) {
                    if (i < PrepCount<D>::N3)
                        prep3<D>(W.g, W.a, i, W.F);
                    else
                        prep4<D>(i - PrepCount<D>::N3, W.F);
                }
                wave_sync();
                // entries 1 .. NV on lanes 0 .. NV - 1 and entry 0 (the
                // intensity gradient, the cheapest) as lane 0's second item
                // when NV = 65 > 64
                for (int q = lane + 1; q <= NV; q += 64) {
                    const int e = q == NV ? 0 : q;
                    int i = e, l = e;
                    if (e >= P) upper_ij(e - P, P, i, l);
                    xput(prec + (size_t)a * NVS + e, (e < P) ? fin_grad<D>(W.F, e) : fin_hess<D>(W.F, i, l));
                }
                wave_sync();  // W is reused by the wave's next angle
            }
        }
        GT_FIT_TICK(1);
        GT_FIT_MEET();
        // sum the per-angle records over all angles in one fixed order (angles
        // a, a + 64, a + 128, a + 192 on lane a, then the wave's DPP sum); all
        // loads first
        {
            constexpr int NE = (NV + FIT_WAVES - 1) / FIT_WAVES;
            double v[NE];
#pragma unroll
            for (int q = 0; q < NE; q++) {
                const int e = wave + q * FIT_WAVES;
                double t = 0.0;
                if (e < NV) {
#pragma unroll
                    for (int u = 0; u < NAMAX / 64; u++)
                        if (lane + 64 * u < nA) t += xget(prec + (size_t)(lane + 64 * u) * NVS + e);
                }
                v[q] = t;
            }
#pragma unroll
            for (int q = 0; q < NE; q++) {
                const int e = wave + q * FIT_WAVES;
                const double t = wave_sum_f64(v[q]);
                if (lane == 0 && e < NV) {
                    if (e < P) {
                        L.g[e] = t;
                    } else {
                        int i, l;
                        upper_ij(e - P, P, i, l);
                        L.H[i * P + l] = t;
                        L.H[l * P + i] = t;
                    }
                }
            }
        }
        __syncthreads();
        GT_FIT_TICK(2);
        // ---- phase D: damped Newton system, solve, projections ------------
        // h is read here, between the two barriers of this update: other
        // members write its new value only after the second one
        const int rk = S.reg_kind;
        if (tid == 0) {
            L.hj = S.h[(size_t)L.c.blk * N + j];
            L.rv_old = rk != GT_REG_NULL ? reg_eval<D>(rk, S.reg_p, iso, L.th_old, L.rg, L.rH) : 0.0;
        }
        __syncthreads();
        // (H + reg'' + I/h) dx = -(g + reg') over the solve's parameters
        // (NewtonGaussian.py:61-74 and :147-152)
        if (tid < PI * PI) {
            const int i = tid / PI, l = tid - (tid / PI) * PI;
            double v = iso ? collapse_H<D>(L.H, P, i, l) : L.H[i * P + l];
            if (rk != GT_REG_NULL) v += L.rH[tid];
            if (i == l) v += 1.0 / L.hj;
            L.A[tid] = v;
        }
        if (tid < PI) {
            double gv;
            if (iso && tid == 1 + D) {
                gv = 0.0;
                for (int u = 0; u < D; u++) gv += L.g[1 + D + u];
            } else {
                gv = L.g[tid];
            }
            if (rk != GT_REG_NULL) gv += L.rg[tid];
            L.b[tid] = -gv;
        }
        __syncthreads();
        if (wave == 0 && S.block == 2) {
            // block gradient descent (GD_lib.py:135-166): block t of the
            // dims order, its gradient g_t (+ the regulariser's), step size
            // h = min(h, 1 / max(|g_t|, 1e-8)), T = old - h g_t, no projection
            const int t = S.gd_blocks[L.c.blk];
            const int o = t == 0 ? 0 : (t == 1 ? 1 : 1 + D), n = t == 0 ? 1 : (t == 1 ? D : PI - 1 - D);
            if (lane == 0) {
                double g2 = 0.0;
                for (int i = 0; i < n; i++) g2 += L.b[o + i] * L.b[o + i];
                const double h = fmin(L.hj, 1.0 / fmax(sqrt(g2), GD_EPS));
                L.hj = h;
                double *th = L.th_new;
                for (int i = 0; i < 1 + D + NRs; i++) th[i] = L.th_old[i];
                for (int i = 0; i < n; i++) th[o + i] = (double)(float)(L.th_old[o + i] + h * L.b[o + i]);
                // norm(old - T) > tol * norm(old) over the block (GD_lib.py:163)
                double dn = 0.0, on = 0.0;
                for (int i = o; i < o + n; i++) {
                    const double d = th[i] - L.th_old[i];
                    dn += d * d;
                    on += L.th_old[i] * L.th_old[i];
                }
                L.moved = sqrt(dn) > S.tol * sqrt(on);
                L.rv_new = rk != GT_REG_NULL ? reg_eval<D>(rk, S.reg_p, iso, th, nullptr, nullptr) : 0.0;
            }
        } else if (wave == 0) {
            // the parameters this update moves: all, or block blk = I | x | r
            int o = 0, n = PI;
            if (S.block) {
                o = L.c.blk == 0 ? 0 : (L.c.blk == 1 ? 1 : 1 + D);
                n = L.c.blk == 0 ? 1 : (L.c.blk == 1 ? D : PI - 1 - D);
                for (int e = lane; e < n * n; e += 64) L.A2[e] = L.A[(o + e / n) * PI + o + e % n];
                if (lane < n) L.b2[lane] = L.b[o + lane];
                wave_sync();
            }
            bool ok;
            if (!S.block) {
                ok = solve_n<D>(n, L.A, L.b, L.x, lane);
            } else if (n == 1 && L.c.blk == 0) {
                // dx = -df[0] / ddf[0, 0] if ddf[0, 0] > 1e-8 (NewtonGaussian.py:154-158)
                ok = L.A2[0] > 1e-8;
                if (lane == 0 && ok) L.x[0] = L.b2[0] / L.A2[0];
            } else {
                ok = solve_n<D>(n, L.A2, L.b2, L.x, lane);
            }
            if (lane == 0) {
                if (!ok) {
                    if (S.block) {
                        L.skip = 1;  // `break` to the next block
                    } else {
                        L.c.done = 2;
                        L.stop = 1;
                    }
                } else {
                    const double *b = L.x;
                    double *t = L.th_new;  // LDS: indexed by the block's runtime offset
                    for (int i = 0; i < 1 + D + NRs; i++) t[i] = L.th_old[i];
                    for (int i = 0; i < n; i++) t[o + i] += b[i];
                    if (!S.block) {
                        double n2 = 0.0;
                        for (int i = 0; i < PI; i++) n2 += b[i] * b[i];
                        L.dxnorm = sqrt(n2);
                    }
                    // projections: GD_lib.py:502-522; block variant :159-187
                    if (!S.block || L.c.blk == 0) t[0] = fmax(0.0, t[0]);
                    if (S.block && L.c.blk == 1)
                        for (int i = 0; i < D; i++) {
                            const double v = t[1 + i];
                            t[1 + i] = v < -0.99 ? -0.99 : (v > 0.99 ? 0.99 : v);
                        }
                    if (!S.block || L.c.blk == 2) {
                        if (iso) {
                            t[1 + D] = fabs(t[1 + D]);
                        } else {
                            double *rr = t + 1 + D;
                            if (D == 2) {
                                if (rr[0] < 0) {
                                    rr[0] = -rr[0];
                                    rr[2] = -rr[2];
                                }
                                if (rr[1] < 0) rr[1] = -rr[1];
                            } else {
                                if (rr[0] < 0) {
                                    rr[0] = -rr[0];
                                    rr[3] = -rr[3];
                                    rr[5] = -rr[5];
                                }
                                if (rr[1] < 0) {
                                    rr[1] = -rr[1];
                                    rr[4] = -rr[4];
                                }
                                if (rr[2] < 0) rr[2] = -rr[2];
                            }
                        }
                    }
                    // rounded to the FP32 atom storage
                    for (int i = 0; i < 1 + D + NRs; i++) t[i] = (double)(float)t[i];
                    if (S.block) {
                        // norm(old - T) > tol * norm(old) over the block (NewtonGaussian.py:214),
                        // T the float32 atom entry: I and x before their projection
                        // (:170-171, :178-180), r after it (:188-199)
                        double dn = 0.0, on = 0.0;
                        for (int i = o; i < o + n; i++) {
                            const double Ti = L.c.blk < 2 ? (double)(float)(L.th_old[i] + b[i - o]) : L.th_new[i];
                            const double d = Ti - L.th_old[i];
                            dn += d * d;
                            on += L.th_old[i] * L.th_old[i];
                        }
                        L.moved = sqrt(dn) > S.tol * sqrt(on);
                    }
                    L.rv_new = rk != GT_REG_NULL ? reg_eval<D>(rk, S.reg_p, iso, L.th_new, nullptr, nullptr) : 0.0;
                }
            }
        }
        __syncthreads();
        GT_FIT_TICK(3);
        if (L.stop) break;
        if (L.skip) {  // block variant: nothing to try for this block
            if (tid == 0) fit_advance(L.c, S, true);
            __syncthreads();
            continue;
        }
        // ---- phase E: detector Gaussians of the proposal -------------------
        for (int k = tid; k < nOwn; k += FIT_THREADS) {
            PairGeom<D> g;
            atom_from_theta<D>(L.th_new, iso, g);
            angle_from_stash<D>(L.geo[k], Gm.lenT, g);
            pair_setup<D>(g, det, nd, S.sorted, L.rec_new[k]);
        }
        __syncthreads();
        GT_FIT_TICK(4);
        // ---- phase F: per-angle FP64 energy changes -----------------------
        for (int k = next_angle(&L.next[1], lane); k < nOwn; k = next_angle(&L.next[1], lane)) {
            const int a = member + G * k;
            gmem<float> *Ra = (gmem<float> *)S.resid + (size_t)a * slab;
            float2 *dt = reinterpret_cast<float2 *>(L.tab[wave]);
            double acc = stage ? pair_delta4<K, 2>(L.rec_old[k], L.rec_new[k], L.p0, L.p1, Ra,
                                                   stage + (size_t)a * slab, n1, lane, dt)
                               : pair_delta4<K, 0>(L.rec_old[k], L.rec_new[k], L.p0, L.p1, Ra, nullptr, n1, lane, dt);
            acc = wave_sum_f64(acc);
            if (lane == 0) xput(dEa + a, acc);
        }
        GT_FIT_MEET();
        GT_FIT_TICK(5);
        // ---- phase G: accept / reject, step size, flags -------------------
        if (wave == 0) {
            double dE = 0.0;
#pragma unroll
            for (int u = 0; u < NAMAX / 64; u++)
                if (lane + 64 * u < nA) dE += xget(dEa + lane + 64 * u);
            dE = wave_sum_f64(dE);
            if (lane == 0) {
                FitCtl &c = L.c;
                // E_new > E_old with E = F + sum of the regulariser terms: with a
                // non-finite term already present E_old = +inf and nothing
                // exceeds it; a non-finite new term exceeds any finite E_old
                const bool bad_old = !isfinite(L.rv_old), bad_new = !isfinite(L.rv_new);
                bool accept;
                if (!isfinite(dE))
                    accept = false;  // a non-finite proposal is rejected
                else if (c.ninf > 0)
                    accept = true;
                else if (bad_new)
                    accept = false;
                else
                    accept = (rk != GT_REG_NULL ? dE + (L.rv_new - L.rv_old) : dE) <= 0.0;
                if (S.Fhist && member == 0) S.Fhist[c.jj + 1] = c.F + dE;
                double h = L.hj;
                if (!accept) {
                    if (S.block == 2) {
                        h /= 2.0;  // block GD: no effect on the BREAK flags (GD_lib.py:157-160)
                    } else {
                        if (h > 2 * FIT_EPS) {
                            c.brk0 = 0;
                            c.brk1 = 0;
                        }
                        h /= 10.0;
                    }
                } else {
                    c.F += dE;
                    if (rk != GT_REG_NULL) {
                        c.Rsum += (bad_new ? 0.0 : L.rv_new) - (bad_old ? 0.0 : L.rv_old);
                        c.ninf += (int)bad_new - (int)bad_old;
                        c.E = c.F + (c.ninf > 0 ? __builtin_huge_val() : c.Rsum);
                    } else {
                        c.E = c.F;
                    }
                    if (S.block ? L.moved : L.dxnorm > S.tol) {
                        c.brk0 = 0;
                        c.brk1 = 0;
                    }
                    h *= S.block == 2 ? 1.1 : (S.block ? 1.3 : 2.0);
                    // every member writes the same values: its own later reads
                    // of atom j then never depend on another member's timing
                    S.I[j] = (float)L.th_new[0];
                    for (int i = 0; i < D; i++) S.x[(size_t)j * D + i] = (float)L.th_new[1 + i];
                    for (int k = 0; k < NRs; k++) S.r[(size_t)j * NRs + k] = (float)L.th_new[1 + D + k];
                    c.accepted++;
                }
                S.h[(size_t)c.blk * N + j] = S.block == 2 ? fmax(h, GD_EPS) : fmin(fmax(h, FIT_EPS), 1.0 / FIT_EPS);
                c.jj++;
                if (S.Ehist && member == 0) S.Ehist[c.jj] = c.E;
                c.updates++;
                L.accept = accept;
            }
        }
        __syncthreads();
        GT_FIT_TICK(6);
        // ---- phase H: commit the residual of the own angles ----------------
        // with a staging slab the energy pass already committed: undo on
        // reject; without one, commit on accept
        if (stage ? !L.accept : L.accept) {
            for (int k = next_angle(&L.next[2], lane); k < nOwn; k = next_angle(&L.next[2], lane)) {
                const int a = member + G * k;
                gmem<float> *Ra = (gmem<float> *)S.resid + (size_t)a * slab;
                if (stage)
                    pair_restore<K>(L.rec_old[k], L.rec_new[k], Ra, stage + (size_t)a * slab, n1, lane);
                else
                    pair_delta4<K, 1>(L.rec_old[k], L.rec_new[k], L.p0, L.p1, Ra, nullptr, n1, lane,
                                      reinterpret_cast<float2 *>(L.tab[wave]));
            }
        }
        // ---- loop control (NewtonGaussian.py:50-98, :131-222) ---------------
        if (tid == 0) fit_advance(L.c, S, false);
        __syncthreads();
        GT_FIT_TICK(7);
    }
#undef GT_FIT_MEET
#undef GT_FIT_TICK
    if (tid == 0) {
        const bool failed = L.abort_ || (G > 1 && xgetu(bar + 2) != 0u);
        if (member == 0) {
            const FitCtl &c = L.c;
            S.it = c.it;
            S.pos = c.pos;
            S.inner = c.inner;
            S.jj = c.jj;
            S.brk0 = c.brk0;
            S.brk1 = c.brk1;
            S.E = c.E;
            S.updates = c.updates;
            S.accepted = c.accepted;
            S.blk = c.blk;
            S.ninf = c.ninf;
            S.F = c.F;
            S.Rsum = c.Rsum;
            if (timing)
                for (int k = 0; k < 8; k++) S.phase_ticks[k] += ticks[k];
        }
        // done only grows (0 -> 1/2 -> 3), whichever member writes last
        if (member == 0 || failed) atomicMax(&S.done, failed ? 3 : L.c.done);
    }
}

// Fit workgroups the current device holds at once, queried once per device.
// The per-device cache is written under a mutex: host threads may launch
// fits concurrently.
int max_blocks_cached()
{
    static std::mutex mu;
    static int cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return -1;
    std::lock_guard<std::mutex> lock(mu);
    if (cache[dev] > 0) return cache[dev];
    int cus = 0, per3 = 0, per2 = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per3, fit_kernel<3>, FIT_THREADS, 0) != hipSuccess) return -1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per2, fit_kernel<2>, FIT_THREADS, 0) != hipSuccess) return -1;
    cache[dev] = cus * min(per3, per2);
    return cache[dev];
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_fit_scratch_bytes(int nA)
{
    if (nA < 0 || nA > NAMAX) return 0;
    return fit_scratch_size(nA);
}

extern "C" int gt_fit_max_blocks(void)
{
    const int n = max_blocks_cached();
    GT_CHECK_ARG(n > 0, "gt_fit_max_blocks: cannot query the device");
    return n;
}

static int launch_fit(gt_fit_slot *slots, int nslots, const gt_fit_block *blocks, int nblocks, int dim, int budget,
                      void *stream)
{
    hipStream_t st = (hipStream_t)stream;
    prof_begin(st, GT_K_FIT);
    if (dim == 3)
        hipLaunchKernelGGL(fit_kernel<3>, dim3(nblocks), dim3(FIT_THREADS), 0, st, slots, nslots, blocks, budget);
    else
        hipLaunchKernelGGL(fit_kernel<2>, dim3(nblocks), dim3(FIT_THREADS), 0, st, slots, nslots, blocks, budget);
    prof_end(st, GT_K_FIT);
    GT_HIP(hipGetLastError());
    return 0;
}

extern "C" int gt_fit_run(gt_fit_slot *slots, int nslots, int dim, int budget, void *stream)
{
    GT_CHECK_ARG(slots != nullptr || nslots == 0, "gt_fit_run: slots is NULL");
    GT_CHECK_ARG(nslots >= 0 && nslots < (1 << 24), "gt_fit_run: bad slot count %d", nslots);
    GT_CHECK_ARG(dim == 2 || dim == 3, "gt_fit_run: dim must be 2 or 3");
    GT_CHECK_ARG(budget >= 1, "gt_fit_run: budget must be >= 1");
    if (nslots == 0) return 0;
    return launch_fit(slots, nslots, nullptr, nslots, dim, budget, stream);
}

extern "C" int gt_fit_run_blocks(gt_fit_slot *slots, int nslots, const gt_fit_block *blocks, int nblocks, int dim,
                                 int budget, void *stream)
{
    GT_CHECK_ARG(slots != nullptr || nslots == 0, "gt_fit_run_blocks: slots is NULL");
    GT_CHECK_ARG(blocks != nullptr || nblocks == 0, "gt_fit_run_blocks: blocks is NULL");
    GT_CHECK_ARG(nslots >= 0 && nslots < (1 << 24), "gt_fit_run_blocks: bad slot count %d", nslots);
    GT_CHECK_ARG(dim == 2 || dim == 3, "gt_fit_run_blocks: dim must be 2 or 3");
    GT_CHECK_ARG(budget >= 1, "gt_fit_run_blocks: budget must be >= 1");
    if (nblocks == 0 || nslots == 0) return 0;
    const int cap = max_blocks_cached();
    GT_CHECK_ARG(cap > 0, "gt_fit_run_blocks: cannot query the device");
    GT_CHECK_ARG(nblocks > 0 && nblocks <= cap,
                 "gt_fit_run_blocks: %d blocks requested, the device holds %d fit workgroups at once", nblocks, cap);
    return launch_fit(slots, nslots, blocks, nblocks, dim, budget, stream);
}
