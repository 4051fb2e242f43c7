// gt_fit.hip -- device-resident damped-Newton fitter (one workgroup per fit).
//
// Restates NewtonGaussian.linesearch (NewtonGaussian.py:18-109) with
// GD_lib._step (GD_lib.py:492-536) for the null regulariser and the l2
// fidelity.  For each atom j of the outer iteration's permutation:
//
//   1. C = data + R_j - R is never formed: the L2 derivatives of atom j
//      against C equal the derivatives of 1/2||R - data||^2 w.r.t. atom j,
//      so the pixel moments are taken against the stored residual R - data.
//   2. (g, H) of atom j: per-angle FP32 pixel moments + FP64 contraction.
//   3. dx = solve(H + I/h_j, -g) in FP64 with partial pivoting (numpy.linalg.
//      solve, NewtonGaussian.py:73-76); singular -> the fit stops (done = 2).
//   4. _step projections: I >= 0, rows of r with a negative diagonal negated.
//   5. Energy change of the committed FP32 residual, accumulated in FP64 over
//      the union of the old and new footprints; reject iff it is > 0 (or not
//      finite), the reference's E[jj] > E[jj-1] test on the full energies.
//   6. Accepted: atom written back, residual updated over the same pixels.
//      Step size h_j x2 / /10 clamped to [1e-4, 1e4], convergence flags as in
//      NewtonGaussian.py:78-98.
//
// Everything for one slot happens inside one workgroup of 16 waves: no host
// round trip and no inter-workgroup synchronisation; workgroup barriers order
// the phases.  Pixel phases give one (atom, angle) pair to each wave with four
// pixels in flight per lane; the FP64 contraction stages per-pair ingredients
// in LDS (pair_prepare) and spreads the 10 + 55 gradient/Hessian entries over
// the waves (fin_grad / fin_hess, one pair per lane); the solve runs on one
// wave with one matrix row per lane.
#include "gt_pixel.h"

namespace gt {

constexpr int FIT_THREADS = 1024;
constexpr int FIT_WAVES = FIT_THREADS / 64;
constexpr int NAMAX = GT_FIT_MAX_ANGLES;
constexpr int DETMAX = GT_FIT_MAX_DETECTOR;
constexpr int FCH = 64;  // pairs per contraction chunk
constexpr double FIT_EPS = 1e-4;  // NewtonGaussian.py:21

template <int D> struct FitLDS {
    static constexpr int P = Dims<D>::P;
    PairRec rec_old[NAMAX];
    PairRec rec_new[NAMAX];
    PairFin<D> fin[FCH];
    PairAlg<D> alg[FCH];
    float p0[DETMAX], p1[DETMAX];
    double dE[FIT_WAVES];
    double th_old[10], th_new[10];
    double g[10], H[100];
    double A[100], b[10], x[10];
    double dxnorm;
    int accept, j, stop, solved;
};

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

// Energy change (apply == 0) or residual update (apply == 1) of replacing
// rec_old by rec_new over the union box; returns the lane's FP64 partial.
// Both passes evaluate the same FP32 expressions, so the committed residual
// is exactly the one whose energy was tested.
template <int K>
__device__ inline double pair_delta4(const PairRec &ro, const PairRec &rn, const float *p0, const float *p1,
                                     float *__restrict__ Ra, int n1, int lane, int apply)
{
    int lo[2], hi[2];
    box_union(ro, rn, K, lo, hi);
    const int W = (K == 2) ? hi[1] - lo[1] + 1 : 1;
    const int H = hi[0] - lo[0] + 1;
    const long total = (W > 0 && H > 0) ? (long)W * H : 0;
    double acc = 0.0;
    if (total == 0) return acc;
    if (K == 2) {
        // column-owning walk (ColChunk): per-column terms of both Gaussians once per lane
        const float kap = (float)KAPPA;
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
            float *col = Ra + i1;
            for (int r0 = cc.ro; r0 < H; r0 += 4 * cc.R) {
                float old[4], pv0[4];
                bool ok[4];
                size_t off[4];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int r = r0 + u * cc.R;
                    ok[u] = r < H;
                    const int i0 = lo[0] + (ok[u] ? r : 0);
                    off[u] = (size_t)i0 * n1;
                    old[u] = col[off[u]];
                    pv0[u] = p0[i0];
                }
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const float d0o = (pv0[u] - ro.c_hi[0]) - ro.c_lo[0];
                    const float d0n = (pv0[u] - rn.c_hi[0]) - rn.c_lo[0];
                    const float qo = fmaf(fmaf(ro.q[0], d0o, rlo), d0o, clo);
                    const float qn = fmaf(fmaf(rn.q[0], d0n, rln), d0n, cln);
                    const float vo = (qo >= kap) ? ro.alpha * __builtin_amdgcn_exp2f(qo) : 0.0f;
                    const float vn = (qn >= kap) ? rn.alpha * __builtin_amdgcn_exp2f(qn) : 0.0f;
                    const float dR = vn - vo;
                    if (ok[u] && dR != 0.0f) {
                        const float nw = old[u] + dR;
                        if (apply)
                            col[off[u]] = nw;
                        else
                            acc += 0.5 * ((double)nw - (double)old[u]) * ((double)nw + (double)old[u]);
                    }
                }
            }
        }
        return acc;
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
            if (ok[u] && dR != 0.0f) {
                const float nw = old[u] + dR;
                if (apply)
                    Ra[idx[u]] = nw;
                else
                    acc += 0.5 * ((double)nw - (double)old[u]) * ((double)nw + (double)old[u]);
            }
        }
        tr.step();
    }
    return acc;
}

// Solve A x = b (n x n, row-major in LDS) with partial pivoting on one wave,
// lane i holding row i.  Pivot = first row of largest magnitude (LAPACK
// i*amax).  Returns false if a pivot is zero.
template <int n> __device__ inline bool wave_solve(const double *A, const double *b, double *x, int lane)
{
    const bool own = lane < n;
    double a[n], bb = own ? b[lane] : 0.0;
#pragma unroll
    for (int c = 0; c < n; c++) a[c] = own ? A[lane * n + c] : 0.0;
    bool ok = true;
#pragma unroll
    for (int k = 0; k < n; k++) {
        double val = (own && lane >= k) ? fabs(a[k]) : -1.0;
        int idx = lane;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(val, o);
            const int oi = __shfl_xor(idx, o);
            if (ov > val || (ov == val && oi < idx)) {
                val = ov;
                idx = oi;
            }
        }
        if (!(val > 0.0)) ok = false;
        const int p = idx;
        if (p != k) {
#pragma unroll
            for (int c = 0; c < n; c++) {
                const double ak = __shfl(a[c], k), ap = __shfl(a[c], p);
                if (lane == k) a[c] = ap;
                if (lane == p) a[c] = ak;
            }
            const double bk = __shfl(bb, k), bp = __shfl(bb, p);
            if (lane == k) bb = bp;
            if (lane == p) bb = bk;
        }
        double pk[n];
#pragma unroll
        for (int c = 0; c < n; c++) pk[c] = __shfl(a[c], k);
        const double pb = __shfl(bb, k);
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
        const double xi = (lane == i) ? s / a[i] : 0.0;
        xs[i] = __shfl(xi, i);
    }
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < n; i++) x[i] = xs[i];
    return ok;
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

template <int D>
__global__ __launch_bounds__(FIT_THREADS) void fit_kernel(gt_fit_slot *__restrict__ slots, int nslots, int budget)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, P = Dm::P, NR = Dm::NR, NM = Dm::NM;
    constexpr int NV = P + P * (P + 1) / 2;
    __shared__ FitLDS<D> L;
    const int sid = blockIdx.x;
    if (sid >= nslots) return;
    gt_fit_slot &S = slots[sid];
    if (!S.active || S.done) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int iso = S.iso;
    const int PI = iso ? 2 + D : P;  // parameters of the solve
    const int NRs = iso ? 1 : NR;
    const gt_geom G = slot_geom(S, D);
    const int N = S.N, nA = S.nA, n1 = S.n1;
    const size_t slab = (size_t)S.n0 * S.n1;
    const float *det[2] = {S.p0, S.p1};
    const int nd[2] = {S.n0, S.n1};
    for (int i = tid; i < S.n0; i += FIT_THREADS) L.p0[i] = S.p0[i];
    if (K == 2)
        for (int i = tid; i < S.n1; i += FIT_THREADS) L.p1[i] = S.p1[i];

    long long ticks[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long last = wall_clock64();
#define GT_FIT_TICK(k)                             \
    if (tid == 0 && S.phase_ticks) {               \
        const long long now_ = wall_clock64();     \
        ticks[k] += now_ - last;                   \
        last = now_;                               \
    }

    const int my_budget = S.quota > 0 ? min(budget, S.quota) : budget;
    for (int step = 0; step < my_budget; step++) {
        // ---- choose the atom, load its parameters -----------------------
        if (tid == 0) {
            L.stop = S.done != 0 || S.it >= S.max_outer || S.pos >= N;
            if (!L.stop) {
                const int j = S.perm[(size_t)S.it * N + S.pos];
                L.stop = (j < 0 || j >= N);
                if (L.stop) S.done = 2;
                L.j = j;
                if (!L.stop) {
                    L.th_old[0] = S.I[j];
                    for (int i = 0; i < D; i++) L.th_old[1 + i] = S.x[(size_t)j * D + i];
                    for (int k = 0; k < NRs; k++) L.th_old[1 + D + k] = S.r[(size_t)j * NRs + k];
                }
            }
        }
        if (tid < NV) {
            if (tid < P)
                L.g[tid] = 0.0;
            else {
                int i, l;
                upper_ij(tid - P, P, i, l);
                L.H[i * P + l] = 0.0;
                L.H[l * P + i] = 0.0;
            }
        }
        __syncthreads();
        if (L.stop) break;
        const int j = L.j;
        // ---- phase A: detector Gaussians of the current atom -------------
        for (int a = tid; a < nA; a += FIT_THREADS) {
            PairGeom<D> g;
            atom_from_theta<D>(L.th_old, iso, g);
            load_angle<D>(G, a, g);
            pair_setup<D>(g, det, nd, S.sorted, L.rec_old[a]);
        }
        __syncthreads();
        GT_FIT_TICK(0);
        // ---- phase B: pixel moments against the residual ------------------
        for (int a = wave; a < nA; a += FIT_WAVES) {
            float acc[MOMW];
#pragma unroll
            for (int v = 0; v < MOMW; v++) acc[v] = 0.0f;
            if (K == 2)
                pair_moments_rows<GT_MODE_RESID>(L.rec_old[a], L.p0, L.p1, S.resid + (size_t)a * slab, n1, lane, acc);
            else
                pair_moments4<K, GT_MODE_RESID>(L.rec_old[a], L.p0, L.p1, S.resid + (size_t)a * slab, n1, lane, acc);
            float mine = 0.0f;
#pragma unroll
            for (int v = 0; v < 30; v++) {
                const float t = wave_sum_f32(acc[v]);
                if (lane == v) mine = t;
            }
            if (lane < MOMW) S.scratch[(size_t)a * MOMW + lane] = mine;
        }
        __syncthreads();
        GT_FIT_TICK(1);
        // ---- phase C: FP64 gradient and Hessian of atom j ----------------
        for (int c0 = 0; c0 < nA; c0 += FCH) {
            const int cnt = min(FCH, nA - c0);
            // pair k of the chunk runs on lane k / 16 of wave k % 16: the FP64
            // prepare work is spread over all SIMDs instead of one wave
            const int k = lane * FIT_WAVES + wave;
            if (k < cnt) {
                const int a = c0 + k;
                const PairRec &rec = L.rec_old[a];
                const bool empty = rec.lo[0] > rec.hi[0] || (K == 2 && rec.lo[1] > rec.hi[1]);
                PairGeom<D> g;
                atom_from_theta<D>(L.th_old, iso, g);
                load_angle<D>(G, a, g);
                double S1[NM], S2[NM];
                const float *m = S.scratch + (size_t)a * MOMW;
                for (int v = 0; v < NM; v++) {
                    S1[v] = empty ? 0.0 : (double)m[v];
                    S2[v] = empty ? 0.0 : (double)m[15 + v];
                }
                const double pstar[2] = {(double)rec.c_hi[0] + (double)rec.c_lo[0],
                                         (double)rec.c_hi[1] + (double)rec.c_lo[1]};
                pair_prepare<D>(g, pstar, S1, S2, 1, L.fin[k], L.alg[k]);
            }
            __syncthreads();
            for (int e = wave; e < NV; e += FIT_WAVES) {
                int i = e, l = e;
                if (e >= P) upper_ij(e - P, P, i, l);
                double v = 0.0;
                if (lane < cnt) v = (e < P) ? fin_grad<D>(L.fin[lane], e) : fin_hess<D>(L.fin[lane], i, l);
                v = wave_sum_f64(v);
                if (lane == 0) {
                    if (e < P) {
                        L.g[e] += v;
                    } else {
                        L.H[i * P + l] += v;
                        if (l != i) L.H[l * P + i] += v;
                    }
                }
            }
            __syncthreads();
        }
        GT_FIT_TICK(2);
        // ---- phase D: damped Newton system, solve, projections ------------
        if (tid < PI * PI) {
            const int i = tid / PI, l = tid - (tid / PI) * PI;
            double v = iso ? collapse_H<D>(L.H, P, i, l) : L.H[i * P + l];
            if (i == l) v += 1.0 / S.h[j];
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
            L.b[tid] = -gv;
        }
        __syncthreads();
        if (wave == 0) {
            bool ok;
            if (PI == P)
                ok = wave_solve<P>(L.A, L.b, L.x, lane);
            else
                ok = wave_solve<2 + D>(L.A, L.b, L.x, lane);
            if (lane == 0) L.solved = ok;
        }
        __syncthreads();
        if (tid == 0) {
            if (!L.solved) {
                S.done = 2;
                L.stop = 1;
            } else {
                const double *b = L.x;
                double n2 = 0.0;
                for (int i = 0; i < PI; i++) n2 += b[i] * b[i];
                L.dxnorm = sqrt(n2);
                // GD_lib._step projections (GD_lib.py:502-522), result rounded to FP32
                double t[10];
                t[0] = fmax(0.0, L.th_old[0] + b[0]);
                for (int i = 0; i < D; i++) t[1 + i] = L.th_old[1 + i] + b[1 + i];
                if (iso) {
                    t[1 + D] = fabs(L.th_old[1 + D] + b[1 + D]);
                } else {
                    for (int k = 0; k < NR; k++) t[1 + D + k] = L.th_old[1 + D + k] + b[1 + D + k];
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
                for (int i = 0; i < 1 + D + NRs; i++) L.th_new[i] = (double)(float)t[i];
            }
        }
        __syncthreads();
        GT_FIT_TICK(3);
        if (L.stop) break;
        // ---- phase E: detector Gaussians of the proposal -------------------
        for (int a = tid; a < nA; a += FIT_THREADS) {
            PairGeom<D> g;
            atom_from_theta<D>(L.th_new, iso, g);
            load_angle<D>(G, a, g);
            pair_setup<D>(g, det, nd, S.sorted, L.rec_new[a]);
        }
        __syncthreads();
        GT_FIT_TICK(4);
        // ---- phase F: FP64 energy change ---------------------------------
        {
            double acc = 0.0;
            for (int a = wave; a < nA; a += FIT_WAVES)
                acc += pair_delta4<K>(L.rec_old[a], L.rec_new[a], L.p0, L.p1, S.resid + (size_t)a * slab, n1, lane, 0);
            acc = wave_sum_f64(acc);
            if (lane == 0) L.dE[wave] = acc;
        }
        __syncthreads();
        GT_FIT_TICK(5);
        // ---- phase G: accept / reject, step size, flags (one thread) -------
        if (tid == 0) {
            double dE = 0.0;
            for (int w = 0; w < FIT_WAVES; w++) dE += L.dE[w];
            const bool accept = dE <= 0.0;  // false for NaN: a non-finite proposal is rejected
            double h = S.h[j];
            if (!accept) {
                if (h > 2 * FIT_EPS) {
                    S.brk0 = 0;
                    S.brk1 = 0;
                }
                h /= 10.0;
            } else {
                S.E += dE;
                if (L.dxnorm > S.tol) {
                    S.brk0 = 0;
                    S.brk1 = 0;
                }
                h *= 2.0;
                S.I[j] = (float)L.th_new[0];
                for (int i = 0; i < D; i++) S.x[(size_t)j * D + i] = (float)L.th_new[1 + i];
                for (int k = 0; k < NRs; k++) S.r[(size_t)j * NRs + k] = (float)L.th_new[1 + D + k];
                S.accepted++;
            }
            S.h[j] = fmin(fmax(h, FIT_EPS), 1.0 / FIT_EPS);
            S.jj++;
            if (S.Ehist) S.Ehist[S.jj] = S.E;
            S.updates++;
            L.accept = accept;
        }
        __syncthreads();
        GT_FIT_TICK(6);
        // ---- phase H: commit the residual ---------------------------------
        if (L.accept) {
            for (int a = wave; a < nA; a += FIT_WAVES)
                pair_delta4<K>(L.rec_old[a], L.rec_new[a], L.p0, L.p1, S.resid + (size_t)a * slab, n1, lane, 1);
        }
        // ---- loop control (NewtonGaussian.py:50-98) ------------------------
        if (tid == 0) {
            S.inner++;
            if (S.brk1 || S.inner >= S.max_inner) {
                S.inner = 0;
                S.pos++;
            }
            S.brk1 = 1;  // BREAK[1] = True at the start of every inner iteration
            if (S.pos >= N) {
                S.pos = 0;
                if (S.brk0 && S.it > S.min_iter) {
                    S.done = 1;
                } else {
                    S.it++;
                    if (S.it >= S.max_outer) S.done = 1;
                    S.brk0 = 1;
                }
            }
        }
        __syncthreads();
        GT_FIT_TICK(7);
    }
#undef GT_FIT_TICK
    if (tid == 0 && S.phase_ticks)
        for (int k = 0; k < 8; k++) S.phase_ticks[k] += ticks[k];
}

}  // namespace gt

using namespace gt;

extern "C" int gt_fit_run(gt_fit_slot *slots, int nslots, int dim, int budget, void *stream)
{
    GT_CHECK_ARG(slots != nullptr || nslots == 0, "gt_fit_run: slots is NULL");
    GT_CHECK_ARG(nslots >= 0 && nslots < (1 << 24), "gt_fit_run: bad slot count %d", nslots);
    GT_CHECK_ARG(dim == 2 || dim == 3, "gt_fit_run: dim must be 2 or 3");
    GT_CHECK_ARG(budget >= 1, "gt_fit_run: budget must be >= 1");
    if (nslots == 0) return 0;
    hipStream_t st = (hipStream_t)stream;
    prof_begin(st, GT_K_FIT);
    if (dim == 3)
        hipLaunchKernelGGL(fit_kernel<3>, dim3(nslots), dim3(FIT_THREADS), 0, st, slots, nslots, budget);
    else
        hipLaunchKernelGGL(fit_kernel<2>, dim3(nslots), dim3(FIT_THREADS), 0, st, slots, nslots, budget);
    prof_end(st, GT_K_FIT);
    GT_HIP(hipGetLastError());
    return 0;
}
