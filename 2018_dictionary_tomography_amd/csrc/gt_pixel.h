// gt_pixel.h -- per-lane pixel loops over one (atom, angle) footprint box.
//
// A wave walks the pixels of a PairRec box in row-major order, 64 pixels per
// step, keeping (row, col) incrementally so no integer division is done per
// pixel.  Pixels with m > 1 contribute nothing (NewtonBases.py:49-56, 375-380);
// the moment walks (moment_pixel) and the 2D-detector forward take the tail
// just outside as it is (below e^-30 of the profile's peak).
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
    __device__ inline void init(int lane, int W, int cb, int maxw = 64)
    {
        Wc = min(maxw, W - cb);
        if (Wc == 32 || Wc == 64) {
            // full chunks: shifts instead of the per-lane integer division
            const int sh = Wc == 32 ? 5 : 6;
            R = 64 >> sh;
            ro = lane >> sh;
            c = lane & (Wc - 1);
        } else {
            R = 64 / Wc;
            ro = lane / Wc;
            c = lane - ro * Wc;
        }
        act = ro < R;
    }
};

// Slack (in units of the exp2 exponent q = kappa m) by which the per-column
// row spans below are widened: far above the FP32 rounding of q (|q| is a
// few tens inside a footprint box), so every pixel whose FP32 q passes the
// cutoff lies inside its span.
constexpr float SPAN_SLACK = 1.0e-2f;

// Step dx of a uniform grid p[i] = p[0] + i dx that every one of the n
// coordinates matches to within dx / 4 (one wave; the value is the same in
// every lane), or 0 if p is not such a grid.
__device__ inline float grid_step(const float *p, int n, int lane)
{
    if (n < 2) return 0.0f;
    const float dx = (p[n - 1] - p[0]) / (float)(n - 1);
    bool ok = dx > 0.0f;
    for (int i = lane; i < n; i += 64) ok = ok && fabsf(p[i] - fmaf((float)i, dx, p[0])) <= 0.25f * dx;
    return __ballot(!ok) == 0ull ? dx : 0.0f;
}

// Row table of one pair over a window of H box rows, in the wave's LDS:
// tab[r] = (x, x^2, x^3, x^4), x = dp0 of row row0 + r (the same FP32
// expressions a pixel would evaluate).  Returns the spacing dx of a uniform
// grid x(r) = x(0) + r dx that every row matches to within dx / 4, or 0 if
// the rows are not such a grid (then no row span is derived from it).
__device__ inline float row_table(const PairRec &rec, const float *p0, int row0, int H, int lane, float4 *tab)
{
    for (int r = lane; r < H; r += 64) {
        const float x = (p0[row0 + r] - rec.c_hi[0]) - rec.c_lo[0];
        const float x2 = x * x;
        tab[r] = make_float4(x, x2, x2 * x, x2 * x2);
    }
    wave_sync();
    if (H < 2) return 0.0f;
    const float x0 = tab[0].x, dx = (tab[H - 1].x - x0) / (float)(H - 1);
    bool ok = dx > 0.0f;
    for (int r = lane; r < H; r += 64) ok = ok && fabsf(tab[r].x - fmaf((float)r, dx, x0)) <= 0.25f * dx;
    return __ballot(!ok) == 0ull ? dx : 0.0f;
}

// Rows [rs, re] (window-relative, inclusive) of one column that can hold
// pixels of the support {q >= kappa}: the roots of the column's quadratic
// q(x) = q0 x^2 + rowlin x + colt (q0 < 0) widened by SPAN_SLACK, mapped to
// rows through the uniform grid and padded by one row.  inv2a = 0.5 / (-q0)
// and inv_dx = 1 / dx are the caller's (approximate reciprocals do: the
// padding covers their error).  Without a grid (dx = 0) or for a flat
// quadratic, the whole window.  rs > re: no support.
__host__ __device__ inline void column_span(const PairRec &rec, float rowlin, float colt, float x0, float dx,
                                            float inv_dx, float inv2a, int H, int &rs, int &re)
{
    rs = 0;
    re = H - 1;
    const float a = -rec.q[0];
    if (!(dx > 0.0f) || !(a > 0.0f)) return;
    const float disc = fmaf(rowlin, rowlin, 4.0f * a * (colt - (float)KAPPA + SPAN_SLACK));
    if (!(disc >= 0.0f)) {
        rs = 1;
        re = 0;
        return;
    }
    const float sq = sqrtf(disc);
    const float ulo = fmaf(rowlin - sq, inv2a, -x0) * inv_dx, uhi = fmaf(rowlin + sq, inv2a, -x0) * inv_dx;
    const float hf = (float)H;
    const int a0 = (int)floorf(fminf(fmaxf(ulo, -2.0f), hf)) - 1, a1 = (int)ceilf(fminf(fmaxf(uhi, -2.0f), hf)) + 1;
    rs = a0 < 0 ? 0 : a0;
    re = a1 > H - 1 ? H - 1 : a1;
}

// The reciprocals column_span takes: 0.5 / (-q0) and 1 / dx (0 when unused).
__host__ __device__ inline float span_inv2a(const PairRec &rec) { return -rec.q[0] > 0.0f ? 0.5f / -rec.q[0] : 0.0f; }
__host__ __device__ inline float span_inv_dx(float dx) { return dx > 0.0f ? 1.0f / dx : 0.0f; }

// Wave minimum / maximum through DPP (no LDS round trips): a running
// min / max over the previous 1, 2, 4, 8 lanes of each 16-lane row (lanes
// before the row start contribute the identity), then lane 15 of each row
// broadcast into the next row (rows 1 and 3) and lane 31 into rows 2 and 3:
// lane 63 ends with the value of the whole wave, which readlane makes
// uniform.  min / max are idempotent, so the overlapping windows are exact.
__device__ inline int wave_min_i32(int v)
{
    constexpr int ID = 0x7fffffff;
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = min(v, __builtin_amdgcn_update_dpp(ID, v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ inline int wave_max_i32(int v)
{
    constexpr int ID = -0x7fffffff - 1;
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x111, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x112, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x114, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x118, 0xf, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x142, 0xa, 0xf, false));
    v = max(v, __builtin_amdgcn_update_dpp(ID, v, 0x143, 0xc, 0xf, false));
    return __builtin_amdgcn_readlane(v, 63);
}

// Can the support {q >= kappa} of a record meet the rectangle [xa, xb] x
// [ya, yb] of detector offsets from its centre?  q is concave with its
// maximum at the centre (the expansion point), so the rectangle is met when
// it holds the centre, else iff the largest value on its four edges (each a
// concave quadratic in one variable, maximised at its clamped vertex)
// reaches the cutoff, widened by SPAN_SLACK.  A record whose quadratic is
// not strictly concave (rounded) is always kept.  f1, f2: the record's
// floats 4..11, (c_lo1, q0, q1, q2) and (q3, q4, q5, pad).
__host__ __device__ inline bool support_meets_rect(const float4 &f1, const float4 &f2, float xa, float xb, float ya, float yb)
{
    const float q0 = f1.y, q1 = f1.z, q2 = f1.w, q3 = f2.x, q4 = f2.y, q5 = f2.z;
    if (!(q0 < 0.0f) || !(q2 < 0.0f) || !(q1 * q1 < 4.0f * q0 * q2)) return true;
    if (xa <= 0.0f && xb >= 0.0f && ya <= 0.0f && yb >= 0.0f) return true;
    const float i0 = -0.5f / q0, i2 = -0.5f / q2;
    float best = -3.0e38f;
#pragma unroll
    for (int e = 0; e < 2; e++) {
        // edges x = const: concave in y
        const float x = e ? xb : xa;
        const float lin = fmaf(q1, x, q4);
        const float y = fminf(fmaxf(lin * i2, ya), yb);
        best = fmaxf(best, fmaf(fmaf(q2, y, lin), y, fmaf(fmaf(q0, x, q3), x, q5)));
        // edges y = const: concave in x
        const float yy = e ? yb : ya;
        const float lin2 = fmaf(q1, yy, q3);
        const float xx = fminf(fmaxf(lin2 * i0, xa), xb);
        best = fmaxf(best, fmaf(fmaf(q0, xx, lin2), xx, fmaf(fmaf(q2, yy, q4), yy, q5)));
    }
    return best >= (float)KAPPA - SPAN_SLACK;
}

// One pixel of the moments walks, branch-free.  xp = (x, x^2, x^3, x^4).
// The walks visit a few pixels just outside the support {m <= 1} (the rows of
// a column chunk's span hull, or of the whole box): there E = 2^q is below
// 2^kappa = e^-30 of the profile's peak and falls off as a Gaussian, so the
// moment sums take it as it is instead of testing every pixel against the
// cut-off; what that adds is below 1e-13 of the sums, under their FP32
// resolution.  The L2 energy terms, which do not vanish outside the support,
// are taken for pixels inside it only.
template <int MODE>
__device__ inline void moment_pixel(const PairRec &rec, float cv, float4 xp, float rowlin, float colt, float *A,
                                    float *B, float *acc)
{
    const float x = xp.x;
    const float q = fmaf(fmaf(rec.q[0], x, rowlin), x, colt);
    const float E = __builtin_amdgcn_exp2f(q);
    float w1;
    if (MODE == GT_MODE_L2) {
        const bool in = q >= (float)KAPPA;
        const float res = fmaf(rec.alpha, E, -cv);
        w1 = res * E;
        acc[30] = fmaf(in ? res : 0.0f, res, acc[30]);
        acc[31] = fmaf(in ? cv : 0.0f, cv, acc[31]);
    } else {
        w1 = cv * E;
    }
    A[0] += w1;
    A[1] = fmaf(w1, xp.x, A[1]);
    A[2] = fmaf(w1, xp.y, A[2]);
    A[3] = fmaf(w1, xp.z, A[3]);
    A[4] = fmaf(w1, xp.w, A[4]);
    if (MODE != GT_MODE_LINEAR) {
        const float w2 = E * E;
        B[0] += w2;
        B[1] = fmaf(w2, xp.x, B[1]);
        B[2] = fmaf(w2, xp.y, B[2]);
        B[3] = fmaf(w2, xp.z, B[3]);
        B[4] = fmaf(w2, xp.w, B[4]);
    }
}

// Rows r0, r0 + step, ... <= he of one lane's column in the moments walk
// (step = RS when RS > 0, a compile-time row-table stride).  The lane's byte
// offset of its row and its row-table entry advance once per group of UNR
// rows; the rows of a group sit at wave-uniform distances (u step rows) from
// them, which the residual loads take as their scalar offsets.
template <int MODE, int UNR, int RS>
__device__ inline void moment_rows(const PairRec &rec, __amdgpu_buffer_rsrc_t crs, unsigned off, unsigned rowb,
                                   const float4 *trow, int step_, int r0, int he, float rowlin, float colt, float *A,
                                   float *B, float *acc)
{
    const int step = RS > 0 ? RS : step_;
    const unsigned gstep = (unsigned)step * rowb;
    // full groups of UNR rows, all inside the hull
    const int last = he - (UNR - 1) * step;
    for (; r0 <= last; r0 += UNR * step) {
        float cv[UNR];
        float4 xp[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            cv[u] = buffer_load_f32(crs, off, (unsigned)u * gstep);
            xp[u] = trow[u * step];
        }
        off += UNR * gstep;
        trow += UNR * step;
#pragma unroll
        for (int u = 0; u < UNR; u++) moment_pixel<MODE>(rec, cv[u], xp[u], rowlin, colt, A, B, acc);
    }
    // remaining rows one at a time (same row order)
    for (; r0 <= he; r0 += step) {
        const float cv = buffer_load_f32(crs, off, 0u);
        moment_pixel<MODE>(rec, cv, *trow, rowlin, colt, A, B, acc);
        off += gstep;
        trow += step;
    }
}

// Widest column chunk of the moments walk: with at least two row phases per
// column, lanes of short and long columns of an ellipse even out better than
// with one lane per column of a 64-wide chunk (measured on the footprints of
// the 3D workloads: 1.44 against 1.55 lane-steps per in-support pixel at
// 128^2, epilogues included).
constexpr int MOM_CHUNK = 32;

// Moments of one 2D-detector pair, separable form: per lane and column,
// A_i = sum_rows w1 dp0^i and B_i = sum_rows E^2 dp0^i (i <= 4); then
// S1[i, j] += A_i dp1^j, S2[i, j] += B_i dp1^j.
//
// The box is taken in windows of at most TROWS rows.  Per window the wave
// writes the row table (tab, TROWS entries of its LDS) and, on a uniform
// detector, the lanes of a column chunk walk only the rows of the hull of
// their columns' spans (column_span), each lane its own rows ro, ro + R, ...
// of it: the pixels left out are outside the support, where the full walk
// adds exact zeros, so the sums are those of walking the whole box.  UNR rows are loaded per step (more rows in flight
// hide load latency when few waves share a SIMD).
template <int MODE, int UNR, int TROWS>
__device__ inline void pair_moments_rows(const PairRec &rec, const float *p0, const float *p1,
                                         const float *__restrict__ Ca, int n1, int lane, float *acc, float4 *tab)
{
    const int W = rec.hi[1] - rec.lo[1] + 1;
    const int Hbox = rec.hi[0] - rec.lo[0] + 1;
    if (W <= 0 || Hbox <= 0) return;
    // residual rows through a raw buffer resource: 32-bit byte offsets (the
    // detector slabs are far below 4 GiB) split into a per-lane part and a
    // wave-uniform part that the load instruction adds itself
    const __amdgpu_buffer_rsrc_t crs = buffer_rsrc(Ca);
    const unsigned rowb = 4u * (unsigned)n1;
    for (int w0 = 0; w0 < Hbox; w0 += TROWS) {
        const int H = min(TROWS, Hbox - w0);
        const int row0 = rec.lo[0] + w0;
        if (w0 > 0) wave_sync();  // the previous window's table is no longer read
        const float dx = row_table(rec, p0, row0, H, lane, tab);
        const float x0 = tab[0].x, inv_dx = span_inv_dx(dx), inv2a = span_inv2a(rec);
        for (int cb = 0; cb < W; cb += MOM_CHUNK) {
            ColChunk cc;
            cc.init(lane, W, cb, MOM_CHUNK);
            const int i1 = rec.lo[1] + cb + (cc.act ? cc.c : 0);
            const float dp1 = (p1[i1] - rec.c_hi[1]) - rec.c_lo[1];
            const float rowlin = fmaf(rec.q[1], dp1, rec.q[3]);
            const float colt = fmaf(fmaf(rec.q[2], dp1, rec.q[4]), dp1, rec.q[5]);
            // rows [hs, he]: the hull of the chunk's column spans; every lane
            // walks its own rows hs + ro, hs + ro + R, ... of it
            int rs, re;
            column_span(rec, rowlin, colt, x0, dx, inv_dx, inv2a, H, rs, re);
            const bool any = cc.act && rs <= re;
            // (clamped to the window: every lane's span lies inside it)
            const int hs = max(wave_min_i32(any ? rs : H), 0), he = min(wave_max_i32(any ? re : -1), H - 1);
            if (he < hs || !cc.act) continue;
            float A[5] = {0, 0, 0, 0, 0}, B[5] = {0, 0, 0, 0, 0};
            const int r0 = hs + cc.ro;
            const unsigned off = 4u * ((unsigned)(row0 + r0) * (unsigned)n1 + (unsigned)i1);
            // two row phases (chunks of 32 columns, the common case): the
            // row-table reads of a group take immediate offsets
            if (cc.R == 2)
                moment_rows<MODE, UNR, 2>(rec, crs, off, rowb, tab + r0, 2, r0, he, rowlin, colt, A, B, acc);
            else
                moment_rows<MODE, UNR, 0>(rec, crs, off, rowb, tab + r0, cc.R, r0, he, rowlin, colt, A, B, acc);
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
    wave_sync();  // tab may be rewritten by the caller's next use
}

// The same moments over the whole box, without spans: BOXW-column chunks,
// each lane its rows ro, ro + R, ... of the box, in windows of at most TROWS
// rows whose x-powers the wave first writes to its row table (tab, the
// expressions of row_table).  For a wave that walks many small boxes whose
// residual rows sit in HBM (the fitter's one atom at a time over every
// angle), the span set-up costs more than the corners it skips.
template <int MODE, int UNR, int TROWS, int BOXW = 64>
__device__ inline void pair_moments_box(const PairRec &rec, const float *p0, const float *p1,
                                        const gmem<float> *__restrict__ Ca, int n1, int lane, float *acc,
                                        float4 *tab)
{
    const int W = rec.hi[1] - rec.lo[1] + 1;
    const int Hbox = rec.hi[0] - rec.lo[0] + 1;
    if (W <= 0 || Hbox <= 0) return;
    const __amdgpu_buffer_rsrc_t crs = buffer_rsrc((const float *)Ca);
    const unsigned rowb = 4u * (unsigned)n1;
    for (int w0 = 0; w0 < Hbox; w0 += TROWS) {
        const int H = min(TROWS, Hbox - w0);
        const int row0 = rec.lo[0] + w0;
        if (w0 > 0) wave_sync();  // the previous window's table is no longer read
        for (int r = lane; r < H; r += 64) {
            const float x = (p0[row0 + r] - rec.c_hi[0]) - rec.c_lo[0];
            const float x2 = x * x;
            tab[r] = make_float4(x, x2, x2 * x, x2 * x2);
        }
        wave_sync();
        for (int cb = 0; cb < W; cb += BOXW) {
            ColChunk cc;
            cc.init(lane, W, cb, BOXW);
            if (!cc.act) continue;
            const int i1 = rec.lo[1] + cb + cc.c;
            const float dp1 = (p1[i1] - rec.c_hi[1]) - rec.c_lo[1];
            const float rowlin = fmaf(rec.q[1], dp1, rec.q[3]);
            const float colt = fmaf(fmaf(rec.q[2], dp1, rec.q[4]), dp1, rec.q[5]);
            float A[5] = {0, 0, 0, 0, 0}, B[5] = {0, 0, 0, 0, 0};
            const int r0 = cc.ro;
            const unsigned off = 4u * ((unsigned)(row0 + r0) * (unsigned)n1 + (unsigned)i1);
            if (cc.R == 2)
                moment_rows<MODE, UNR, 2>(rec, crs, off, rowb, tab + r0, 2, r0, H - 1, rowlin, colt, A, B, acc);
            else
                moment_rows<MODE, UNR, 0>(rec, crs, off, rowb, tab + r0, cc.R, r0, H - 1, rowlin, colt, A, B, acc);
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
    wave_sync();  // tab may be rewritten by the caller's next use
}

// Wave64 sums through DPP row shifts and row broadcasts (VALU only, no LDS
// permutes): an inclusive scan within each row of 16 lanes, then rows 0-1
// and 2-3 joined, then the halves; lane 63 ends with the total, which every
// lane reads back.  The order of the additions is fixed.  Lanes whose DPP
// source lies outside the row (or whose row is masked) add 0.
template <int CTRL, int ROWS> __device__ inline float dpp_f32(float x)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS> __device__ inline double dpp_f64(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b & 0xffffffffll), CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

__device__ inline float wave_sum_f32(float v)
{
    v += dpp_f32<0x111, 0xf>(v);  // row_shr:1
    v += dpp_f32<0x112, 0xf>(v);  // row_shr:2
    v += dpp_f32<0x114, 0xf>(v);  // row_shr:4
    v += dpp_f32<0x118, 0xf>(v);  // row_shr:8
    v += dpp_f32<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
    v += dpp_f32<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

__device__ inline double wave_sum_f64(double v)
{
    v += dpp_f64<0x111, 0xf>(v);
    v += dpp_f64<0x112, 0xf>(v);
    v += dpp_f64<0x114, 0xf>(v);
    v += dpp_f64<0x118, 0xf>(v);
    v += dpp_f64<0x142, 0xa>(v);
    v += dpp_f64<0x143, 0xc>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffll), 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

}  // namespace gt
