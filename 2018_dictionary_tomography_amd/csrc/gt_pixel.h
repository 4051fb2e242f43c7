// gt_pixel.h -- per-lane pixel loops over one (atom, angle) footprint box.
//
// A wave walks the pixels of a PairRec box in row-major order, 64 pixels per
// step, keeping (row, col) incrementally so no integer division is done per
// pixel.  Pixels with m > 1 contribute nothing (NewtonBases.py:49-56, 375-380).
#pragma once

#include "gt_common.h"

namespace gt {

constexpr int MOMW = 32;  // floats per moment record

// Moment accumulation over one pair box.
//   MODE 0 (linear): w1 = C E          acc[0..NM)
//   MODE 1 (L2):     w1 = (aE - C) E, w2 = E^2, acc[30] += (aE-C)^2, acc[31] += C^2
//   MODE 2 (resid):  w1 = C E, w2 = E^2
// acc[15 + m] holds the w2 moments.  Ca points at this angle's [n0][n1] slab.
template <int K, int MODE>
__device__ inline void pair_moments(const PairRec &rec, const float *__restrict__ p0, const float *__restrict__ p1,
                                    const float *__restrict__ Ca, int n1, int lane, float *acc)
{
    const int W = (K == 2) ? rec.hi[1] - rec.lo[1] + 1 : 1;
    const int H = rec.hi[0] - rec.lo[0] + 1;
    const long total = (W > 0 && H > 0) ? (long)W * H : 0;
    if (total == 0) return;
    int r = lane / W, cidx = lane - (lane / W) * W;
    const int dr = 64 / W, dc = 64 - dr * W;
    const float kap = (float)KAPPA;
    for (long k = lane; k < total; k += 64) {
        const int i0 = rec.lo[0] + r;
        const int i1 = (K == 2) ? rec.lo[1] + cidx : 0;
        const float dp0 = (p0[i0] - rec.c_hi[0]) - rec.c_lo[0];
        const float dp1 = (K == 2) ? (p1[i1] - rec.c_hi[1]) - rec.c_lo[1] : 0.0f;
        const float q = pair_q(rec.q, dp0, dp1);
        if (q >= kap) {
            const float E = __builtin_amdgcn_exp2f(q);
            const float cv = Ca[(size_t)i0 * n1 + i1];
            float w1;
            if (MODE == GT_MODE_L2) {
                const float res = fmaf(rec.alpha, E, -cv);
                w1 = res * E;
                acc[30] = fmaf(res, res, acc[30]);
                acc[31] = fmaf(cv, cv, acc[31]);
            } else {
                w1 = cv * E;
            }
            const float w2 = E * E;
            if (K == 2) {
                const float x1 = dp0, y1 = dp1;
                const float x2 = x1 * x1, xy = x1 * y1, y2 = y1 * y1;
                const float mono[15] = {1.0f,    x1,      y1,      x2,      xy,      y2,      x2 * x1, x2 * y1,
                                        x1 * y2, y2 * y1, x2 * x2, x2 * xy, x2 * y2, xy * y2, y2 * y2};
#pragma unroll
                for (int m = 0; m < 15; m++) acc[m] = fmaf(w1, mono[m], acc[m]);
                if (MODE != GT_MODE_LINEAR) {
#pragma unroll
                    for (int m = 0; m < 15; m++) acc[15 + m] = fmaf(w2, mono[m], acc[15 + m]);
                }
            } else {
                const float x1 = dp0, x2 = x1 * x1;
                const float mono[5] = {1.0f, x1, x2, x2 * x1, x2 * x2};
#pragma unroll
                for (int m = 0; m < 5; m++) acc[m] = fmaf(w1, mono[m], acc[m]);
                if (MODE != GT_MODE_LINEAR) {
#pragma unroll
                    for (int m = 0; m < 5; m++) acc[15 + m] = fmaf(w2, mono[m], acc[15 + m]);
                }
            }
        }
        cidx += dc;
        r += dr;
        if (cidx >= W) {
            cidx -= W;
            r++;
        }
    }
}

// Four pixel trackers of a wave over a W-wide box: tracker u visits the
// linear indices lane + 64 u + 256 t, t = 0, 1, ...
struct Track4 {
    int r[4], c[4];
    int dr, dc, W;
    __device__ inline void init(int lane, int W_)
    {
        W = W_;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int k = lane + 64 * u;
            r[u] = k / W;
            c[u] = k - r[u] * W;
        }
        dr = 256 / W;
        dc = 256 - dr * W;
    }
    __device__ inline void step()
    {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            c[u] += dc;
            r[u] += dr;
            if (c[u] >= W) {
                c[u] -= W;
                r[u]++;
            }
        }
    }
};

// pair_moments with four independent pixels per lane per iteration: the
// four loads of C are issued before any of them is used, so one wave keeps
// four memory requests in flight (the fitter runs few waves per CU).
template <int K, int MODE>
__device__ inline void pair_moments4(const PairRec &rec, const float *p0, const float *p1,
                                     const float *__restrict__ Ca, int n1, int lane, float *acc)
{
    const int W = (K == 2) ? rec.hi[1] - rec.lo[1] + 1 : 1;
    const int H = rec.hi[0] - rec.lo[0] + 1;
    const long total = (W > 0 && H > 0) ? (long)W * H : 0;
    if (total == 0) return;
    Track4 tr;
    tr.init(lane, W);
    const float kap = (float)KAPPA;
    for (long k = lane; k < total; k += 256) {
        float cv[4], dp0[4], dp1[4];
        bool ok[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            ok[u] = k + 64 * u < total;
            const int i0 = rec.lo[0] + (ok[u] ? tr.r[u] : 0);
            const int i1 = (K == 2) ? rec.lo[1] + (ok[u] ? tr.c[u] : 0) : 0;
            cv[u] = Ca[(size_t)i0 * n1 + i1];
            dp0[u] = (p0[i0] - rec.c_hi[0]) - rec.c_lo[0];
            dp1[u] = (K == 2) ? (p1[i1] - rec.c_hi[1]) - rec.c_lo[1] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float q = pair_q(rec.q, dp0[u], dp1[u]);
            if (ok[u] && q >= kap) {
                const float E = __builtin_amdgcn_exp2f(q);
                float w1;
                if (MODE == GT_MODE_L2) {
                    const float res = fmaf(rec.alpha, E, -cv[u]);
                    w1 = res * E;
                    acc[30] = fmaf(res, res, acc[30]);
                    acc[31] = fmaf(cv[u], cv[u], acc[31]);
                } else {
                    w1 = cv[u] * E;
                }
                const float w2 = E * E;
                if (K == 2) {
                    const float x1 = dp0[u], y1 = dp1[u];
                    const float x2 = x1 * x1, xy = x1 * y1, y2 = y1 * y1;
                    const float mono[15] = {1.0f,    x1,      y1,      x2,      xy,      y2,      x2 * x1, x2 * y1,
                                            x1 * y2, y2 * y1, x2 * x2, x2 * xy, x2 * y2, xy * y2, y2 * y2};
#pragma unroll
                    for (int m = 0; m < 15; m++) acc[m] = fmaf(w1, mono[m], acc[m]);
                    if (MODE != GT_MODE_LINEAR) {
#pragma unroll
                        for (int m = 0; m < 15; m++) acc[15 + m] = fmaf(w2, mono[m], acc[15 + m]);
                    }
                } else {
                    const float x1 = dp0[u], x2 = x1 * x1;
                    const float mono[5] = {1.0f, x1, x2, x2 * x1, x2 * x2};
#pragma unroll
                    for (int m = 0; m < 5; m++) acc[m] = fmaf(w1, mono[m], acc[m]);
                    if (MODE != GT_MODE_LINEAR) {
#pragma unroll
                        for (int m = 0; m < 5; m++) acc[15 + m] = fmaf(w2, mono[m], acc[15 + m]);
                    }
                }
            }
        }
        tr.step();
    }
}

// Column-owning layout of a box of width W for one wave: for W <= 64 the wave
// covers R = 64 / W rows at a time (lane = ro * W + c); wider boxes are walked
// in 64-column chunks with R = 1.  A lane keeps its column for a whole chunk,
// so every per-column quantity is computed once and the row loop only sees
// terms in dp0.  Consecutive lanes read consecutive columns (coalesced).
struct ColChunk {
    int c, ro, R, Wc;
    bool act;
    __device__ inline void init(int lane, int W, int cb)
    {
        Wc = min(64, W - cb);
        R = 64 / Wc;
        ro = lane / Wc;
        c = lane - ro * Wc;
        act = ro < R;
    }
};

// Moments of one 2D-detector pair, separable form: per lane and column,
// A_i = sum_rows w1 dp0^i and B_i = sum_rows E^2 dp0^i (i <= 4); then
// S1[i, j] += A_i dp1^j, S2[i, j] += B_i dp1^j.  Same sums as pair_moments.
// UNR rows are loaded per step (more rows in flight hide load latency when
// few waves share a SIMD).
// One pixel of pair_moments_rows: branch-free cutoff (outside the support
// E = 0 and every accumulation adds an exact zero, so the sums equal those of
// skipping the pixel).
template <int MODE>
__device__ inline void moment_pixel(const PairRec &rec, float cv, float x, float rowlin, float colt, float *A,
                                    float *B, float *acc)
{
    const float q = fmaf(fmaf(rec.q[0], x, rowlin), x, colt);
    const bool in = q >= (float)KAPPA;
    const float E = in ? __builtin_amdgcn_exp2f(q) : 0.0f;
    float w1;
    if (MODE == GT_MODE_L2) {
        const float res = fmaf(rec.alpha, E, -cv);
        w1 = res * E;
        acc[30] = fmaf(in ? res : 0.0f, res, acc[30]);
        acc[31] = fmaf(in ? cv : 0.0f, cv, acc[31]);
    } else {
        w1 = cv * E;
    }
    const float x2 = x * x;
    const float xp[5] = {1.0f, x, x2, x2 * x, x2 * x2};
#pragma unroll
    for (int i = 0; i < 5; i++) A[i] = fmaf(w1, xp[i], A[i]);
    if (MODE != GT_MODE_LINEAR) {
        const float w2 = E * E;
#pragma unroll
        for (int i = 0; i < 5; i++) B[i] = fmaf(w2, xp[i], B[i]);
    }
}

template <int MODE, int UNR = 4>
__device__ inline void pair_moments_rows(const PairRec &rec, const float *p0, const float *p1,
                                         const float *__restrict__ Ca, int n1, int lane, float *acc)
{
    const int W = rec.hi[1] - rec.lo[1] + 1;
    const int H = rec.hi[0] - rec.lo[0] + 1;
    if (W <= 0 || H <= 0) return;
    // 32-bit byte offsets from the wave-uniform bases (scalar-base loads, one
    // multiply-add per pixel); detector slabs are far below 4 GiB
    const char *cbase = reinterpret_cast<const char *>(Ca);
    const char *pbase = reinterpret_cast<const char *>(p0);
    const unsigned rowb = 4u * (unsigned)n1;
    for (int cb = 0; cb < W; cb += 64) {
        ColChunk cc;
        cc.init(lane, W, cb);
        if (!cc.act) continue;
        const int i1 = rec.lo[1] + cb + cc.c;
        const float dp1 = (p1[i1] - rec.c_hi[1]) - rec.c_lo[1];
        const float rowlin = fmaf(rec.q[1], dp1, rec.q[3]);
        const float colt = fmaf(fmaf(rec.q[2], dp1, rec.q[4]), dp1, rec.q[5]);
        float A[5] = {0, 0, 0, 0, 0}, B[5] = {0, 0, 0, 0, 0};
        const unsigned cb0 = 4u * ((unsigned)rec.lo[0] * (unsigned)n1 + (unsigned)i1);
        const unsigned pb0 = 4u * (unsigned)rec.lo[0];
        int r0 = cc.ro;
        // full groups of UNR rows, all inside the box
        for (; r0 + (UNR - 1) * cc.R < H; r0 += UNR * cc.R) {
            float cv[UNR], dp0[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                const unsigned r = (unsigned)(r0 + u * cc.R);
                cv[u] = *reinterpret_cast<const float *>(cbase + (cb0 + r * rowb));
                dp0[u] = (*reinterpret_cast<const float *>(pbase + (pb0 + 4u * r)) - rec.c_hi[0]) - rec.c_lo[0];
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) moment_pixel<MODE>(rec, cv[u], dp0[u], rowlin, colt, A, B, acc);
        }
        // remaining rows one at a time (same row order)
        for (; r0 < H; r0 += cc.R) {
            const unsigned r = (unsigned)r0;
            const float cv = *reinterpret_cast<const float *>(cbase + (cb0 + r * rowb));
            const float x = (*reinterpret_cast<const float *>(pbase + (pb0 + 4u * r)) - rec.c_hi[0]) - rec.c_lo[0];
            moment_pixel<MODE>(rec, cv, x, rowlin, colt, A, B, acc);
        }
        const float y2 = dp1 * dp1;
        const float yp[5] = {1.0f, dp1, y2, y2 * dp1, y2 * y2};
#pragma unroll
        for (int i = 0; i <= 4; i++)
#pragma unroll
            for (int j = 0; i + j <= 4; j++) {
                const int m = (i + j) * (i + j + 1) / 2 + j;
                acc[m] = fmaf(A[i], yp[j], acc[m]);
                if (MODE != GT_MODE_LINEAR) acc[15 + m] = fmaf(B[i], yp[j], acc[15 + m]);
            }
    }
}

__device__ inline float wave_sum_f32(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ inline double wave_sum_f64(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

}  // namespace gt
