// gt_fit.hip -- device-resident damped-Newton fitter (one workgroup per fit).
//
// Restates NewtonGaussian.linesearch (NewtonGaussian.py:18-109) with
// GD_lib._step (GD_lib.py:492-536) for the null regulariser and the l2
// fidelity.  For each atom j of the outer iteration's permutation:
//
//   1. C = data + R_j - R is never formed: the L2 derivatives of atom j
//      against C equal the derivatives of 1/2||R - data||^2 w.r.t. atom j,
//      so the pixel moments are taken against the stored residual R - data.
//   2. (f, g, H) of atom j: per-angle FP32 pixel moments + FP64 contraction.
//   3. dx = solve(H + I/h_j, -g) in FP64 with partial pivoting (numpy.linalg.
//      solve, NewtonGaussian.py:73-76); singular -> the fit stops (done = 2).
//   4. _step projections: I >= 0, rows of r with a negative diagonal negated.
//   5. Energy change of the committed FP32 residual, accumulated in FP64 over
//      the union of the old and new footprints; reject iff it is > 0 (or not
//      finite), exactly the reference's E[jj] > E[jj-1] test on the
//      full energies.
//   6. Accepted: atom written back, residual updated over the same pixels.
//      Step size h_j x2 / /10 clamped to [1e-4, 1e4], convergence flags as in
//      NewtonGaussian.py:78-98.
//
// Everything for one slot happens inside one workgroup (16 waves): no host
// round trip and no inter-workgroup synchronisation.  Between the phases the
// workgroup barrier orders LDS and global-memory traffic of the slot.  The
// pixel phases give one (atom, angle) pair to each wave; the FP64 contraction
// gives one pair to each thread and reduces with wave shuffles.
#include "gt_pixel.h"

namespace gt {

constexpr int FIT_THREADS = 1024;
constexpr int FIT_WAVES = FIT_THREADS / 64;
constexpr int NAMAX = GT_FIT_MAX_ANGLES;
constexpr double FIT_EPS = 1e-4;  // NewtonGaussian.py:21

template <int D> struct FitLDS {
    static constexpr int P = Dims<D>::P;
    static constexpr int NV = P + P * (P + 1) / 2;
    PairRec rec_old[NAMAX];
    PairRec rec_new[NAMAX];
    double wred[FIT_WAVES][NV];
    double dE[FIT_WAVES];
    double th_old[10], th_new[10];
    double g[10], H[100];
    double dxnorm;
    int accept, j, stop;
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
    for (int k = 0; k < 2; k++) {
        const bool ea = a.lo[0] > a.hi[0] || (K == 2 && a.lo[1] > a.hi[1]);
        const bool eb = b.lo[0] > b.hi[0] || (K == 2 && b.lo[1] > b.hi[1]);
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
template <int K>
__device__ inline double pair_delta(const PairRec &ro, const PairRec &rn, const float *__restrict__ p0,
                                    const float *__restrict__ p1, float *__restrict__ Ra, int n1, int lane, int apply)
{
    int lo[2], hi[2];
    box_union(ro, rn, K, lo, hi);
    const int W = (K == 2) ? hi[1] - lo[1] + 1 : 1;
    const int H = hi[0] - lo[0] + 1;
    const long total = (W > 0 && H > 0) ? (long)W * H : 0;
    double acc = 0.0;
    if (total == 0) return acc;
    int r = lane / W, c = lane - (lane / W) * W;
    const int dr = 64 / W, dc = 64 - dr * W;
    for (long k = lane; k < total; k += 64) {
        const int i0 = lo[0] + r;
        const int i1 = (K == 2) ? lo[1] + c : 0;
        const float pv0 = p0[i0];
        const float pv1 = (K == 2) ? p1[i1] : 0.0f;
        const float dR = pair_value(rn, pv0, pv1) - pair_value(ro, pv0, pv1);
        if (dR != 0.0f) {
            float *px = Ra + (size_t)i0 * n1 + i1;
            const float old = *px;
            const float nw = old + dR;
            if (apply) {
                *px = nw;
            } else {
                acc += 0.5 * ((double)nw - (double)old) * ((double)nw + (double)old);
            }
        }
        c += dc;
        r += dr;
        if (c >= W) {
            c -= W;
            r++;
        }
    }
    return acc;
}

// LU solve with partial pivoting, n <= 10; returns false if singular.
__device__ inline bool lu_solve(double *A, double *b, int n)
{
    for (int k = 0; k < n; k++) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i * n + k]) > best) {
                best = fabs(A[i * n + k]);
                p = i;
            }
        if (!(best > 0.0)) return false;
        if (p != k) {
            for (int c = 0; c < n; c++) {
                const double t = A[k * n + c];
                A[k * n + c] = A[p * n + c];
                A[p * n + c] = t;
            }
            const double t = b[k];
            b[k] = b[p];
            b[p] = t;
        }
        for (int i = k + 1; i < n; i++) {
            const double f = A[i * n + k] / A[k * n + k];
            for (int c = k; c < n; c++) A[i * n + c] -= f * A[k * n + c];
            b[i] -= f * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (int c = i + 1; c < n; c++) s -= A[i * n + c] * b[c];
        b[i] = s / A[i * n + i];
    }
    return true;
}

template <int D>
__global__ __launch_bounds__(FIT_THREADS) void fit_kernel(gt_fit_slot *__restrict__ slots, int nslots, int budget)
{
    using Dm = Dims<D>;
    constexpr int K = Dm::K, P = Dm::P, NR = Dm::NR, NM = Dm::NM, NV = FitLDS<D>::NV;
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

    for (int step = 0; step < budget; step++) {
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
        // ---- phase B: pixel moments against the residual ------------------
        for (int a = wave; a < nA; a += FIT_WAVES) {
            float acc[MOMW];
#pragma unroll
            for (int v = 0; v < MOMW; v++) acc[v] = 0.0f;
            pair_moments<K, GT_MODE_RESID>(L.rec_old[a], S.p0, S.p1, S.resid + (size_t)a * slab, n1, lane, acc);
            float mine = 0.0f;
#pragma unroll
            for (int v = 0; v < MOMW; v++) {
                const float t = wave_sum_f32(acc[v]);
                if (lane == v) mine = t;
            }
            if (lane < MOMW) S.scratch[(size_t)a * MOMW + lane] = mine;
        }
        __syncthreads();
        // ---- phase C: FP64 gradient and Hessian of atom j ----------------
        {
            double gl[P], hl[P * P];
            for (int i = 0; i < P; i++) gl[i] = 0.0;
            for (int i = 0; i < P * P; i++) hl[i] = 0.0;
            for (int a = tid; a < nA; a += FIT_THREADS) {
                const PairRec &rec = L.rec_old[a];
                const bool empty = rec.lo[0] > rec.hi[0] || (K == 2 && rec.lo[1] > rec.hi[1]);
                if (empty) continue;
                PairGeom<D> g;
                atom_from_theta<D>(L.th_old, iso, g);
                load_angle<D>(G, a, g);
                double S1[NM], S2[NM];
                const float *m = S.scratch + (size_t)a * MOMW;
                for (int k = 0; k < NM; k++) {
                    S1[k] = m[k];
                    S2[k] = m[15 + k];
                }
                const double pstar[2] = {(double)rec.c_hi[0] + (double)rec.c_lo[0],
                                         (double)rec.c_hi[1] + (double)rec.c_lo[1]};
                pair_finish<D>(g, pstar, S1, S2, 1, 1, gl, hl);
            }
            int k = 0;
            for (int i = 0; i < P; i++) {
                const double v = wave_sum_f64(gl[i]);
                if (lane == 0) L.wred[wave][k] = v;
                k++;
            }
            for (int i = 0; i < P; i++)
                for (int l = i; l < P; l++) {
                    const double v = wave_sum_f64(hl[i * P + l]);
                    if (lane == 0) L.wred[wave][k] = v;
                    k++;
                }
        }
        __syncthreads();
        if (tid < NV) {
            double s = 0.0;
            for (int w = 0; w < FIT_WAVES; w++) s += L.wred[w][tid];
            if (tid < P) {
                L.g[tid] = s;
            } else {
                int k = P;
                for (int i = 0; i < P; i++)
                    for (int l = i; l < P; l++) {
                        if (k == tid) {
                            L.H[i * P + l] = s;
                            L.H[l * P + i] = s;
                        }
                        k++;
                    }
            }
        }
        __syncthreads();
        // ---- phase D: damped Newton step and projections (one thread) -----
        if (tid == 0) {
            double A[100], b[10];
            double gs[10];
            if (iso) {
                // collapse [I, x, r_0..r_{D-1}, offdiag] -> [I, x, rho]
                double Hc[10][10];
                for (int i = 0; i < PI; i++) {
                    for (int l = 0; l < PI; l++) {
                        double s = 0.0;
                        for (int a2 = 0; a2 < P; a2++) {
                            const double ea = (a2 <= D) ? (a2 == i ? 1.0 : 0.0)
                                                        : ((i == 1 + D && a2 < 1 + 2 * D) ? 1.0 : 0.0);
                            if (ea == 0.0) continue;
                            for (int b2 = 0; b2 < P; b2++) {
                                const double eb = (b2 <= D) ? (b2 == l ? 1.0 : 0.0)
                                                            : ((l == 1 + D && b2 < 1 + 2 * D) ? 1.0 : 0.0);
                                if (eb != 0.0) s += L.H[a2 * P + b2];
                            }
                        }
                        Hc[i][l] = s;
                    }
                    double s = 0.0;
                    for (int a2 = 0; a2 < P; a2++) {
                        const double ea = (a2 <= D) ? (a2 == i ? 1.0 : 0.0)
                                                    : ((i == 1 + D && a2 < 1 + 2 * D) ? 1.0 : 0.0);
                        s += ea * L.g[a2];
                    }
                    gs[i] = s;
                }
                for (int i = 0; i < PI; i++)
                    for (int l = 0; l < PI; l++) A[i * PI + l] = Hc[i][l];
            } else {
                for (int i = 0; i < P * P; i++) A[i] = L.H[i];
                for (int i = 0; i < P; i++) gs[i] = L.g[i];
            }
            const double hj = S.h[j];
            for (int i = 0; i < PI; i++) {
                A[i * PI + i] += 1.0 / hj;
                b[i] = -gs[i];
            }
            const bool ok = lu_solve(A, b, PI);
            if (!ok) {
                S.done = 2;
                L.stop = 1;
            } else {
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
                        if (rr[0] < 0) { rr[0] = -rr[0]; rr[2] = -rr[2]; }
                        if (rr[1] < 0) rr[1] = -rr[1];
                    } else {
                        if (rr[0] < 0) { rr[0] = -rr[0]; rr[3] = -rr[3]; rr[5] = -rr[5]; }
                        if (rr[1] < 0) { rr[1] = -rr[1]; rr[4] = -rr[4]; }
                        if (rr[2] < 0) rr[2] = -rr[2];
                    }
                }
                for (int i = 0; i < 1 + D + NRs; i++) L.th_new[i] = (double)(float)t[i];
            }
        }
        __syncthreads();
        if (L.stop) break;
        // ---- phase E: detector Gaussians of the proposal -------------------
        for (int a = tid; a < nA; a += FIT_THREADS) {
            PairGeom<D> g;
            atom_from_theta<D>(L.th_new, iso, g);
            load_angle<D>(G, a, g);
            pair_setup<D>(g, det, nd, S.sorted, L.rec_new[a]);
        }
        __syncthreads();
        // ---- phase F: FP64 energy change ---------------------------------
        {
            double acc = 0.0;
            for (int a = wave; a < nA; a += FIT_WAVES)
                acc += pair_delta<K>(L.rec_old[a], L.rec_new[a], S.p0, S.p1, S.resid + (size_t)a * slab, n1, lane, 0);
            acc = wave_sum_f64(acc);
            if (lane == 0) L.dE[wave] = acc;
        }
        __syncthreads();
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
        // ---- phase H: commit the residual ---------------------------------
        if (L.accept) {
            for (int a = wave; a < nA; a += FIT_WAVES)
                pair_delta<K>(L.rec_old[a], L.rec_new[a], S.p0, S.p1, S.resid + (size_t)a * slab, n1, lane, 1);
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
    }
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
