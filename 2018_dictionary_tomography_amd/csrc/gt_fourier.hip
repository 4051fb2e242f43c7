// gt_fourier.hip -- analytic Fourier transform of Gaussian atoms on a grid.
//
// out(k) = sum_atoms I exp(-|U^T k|^2 / 2) exp(-i x . k)   if |U^T k|^2 < 30
// with (I, x, r) the FT-space parameters (Fourier_Transform.py:293-333); the
// isotropic form uses |U^T k|^2 := |k|^2 / r^2.  This is __grid_eval_aniso /
// __grid_eval_iso, GaussDictCode/bin/FTGaussian3D.py:95-110, 151-164.
//
// Accuracy: the phase x . k reaches hundreds of radians, where an FP32 phase
// would carry ~1e-5 relative error.  The setup stores x / (2 pi) as an FP32
// pair (xh + xl, ~48 bits) and the kernel forms the phase in turns with an
// exact FMA product xh * k = p + e, reduces p exactly by rint, and evaluates
// the hardware sine and cosine (argument in turns) on the small remainder.
// Per-point complex sums are Kahan sums.
#include "gt_common.h"

namespace gt {

struct FtRec {
    float amp;
    float xh[3], xl[3];  // x / (2 pi) as float pairs
    float u[6];          // U^T k coefficients scaled by sqrt(log2(e) / 2)
    float iso;           // 1: isotropic form, u[0] = sqrt(log2(e)/2) / r
    float pad0[2];
    int lo[3], hi[3];
    int pad1[10];
};
static_assert(sizeof(FtRec) == 128, "FtRec must stay 128 bytes");

constexpr double HALF_LOG2E_SQRT = 0.8493218002880191;  // sqrt(log2(e) / 2)
constexpr double FT_CUT = 30.0;                         // |U^T k|^2 < 30
constexpr double TWO_PI = 6.283185307179586;

__global__ __launch_bounds__(256) void ft_setup_kernel(gt_atoms A, gt_kgrid Kg, FtRec *recs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    FtRec rec;
    rec.amp = A.I[j];
    for (int i = 0; i < 3; i++) {
        const double xv = (double)A.x[(size_t)j * 3 + i] / TWO_PI;
        rec.xh[i] = (float)xv;
        rec.xl[i] = (float)(xv - (double)rec.xh[i]);
    }
    for (int k = 0; k < 6; k++) rec.u[k] = 0.0f;
    rec.pad0[0] = rec.pad0[1] = 0.0f;
    for (int k = 0; k < 10; k++) rec.pad1[k] = 0;
    double hw[3];
    bool ok = Kg.sorted != 0;
    if (A.iso) {
        const double r = A.r[j];
        rec.iso = 1.0f;
        rec.u[0] = (float)(HALF_LOG2E_SQRT / r);
        for (int c = 0; c < 3; c++) hw[c] = fabs(r) * sqrt(FT_CUT);
        ok = ok && r != 0.0;
    } else {
        rec.iso = 0.0f;
        double U[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        for (int k = 0; k < 6; k++) {
            U[slot_row(3, k)][slot_col(3, k)] = A.r[(size_t)j * 6 + k];
            rec.u[k] = (float)(HALF_LOG2E_SQRT * A.r[(size_t)j * 6 + k]);
        }
        // {k : |U^T k|^2 < 30}: half-width along k_c = sqrt(30) * |column c of U^{-1}|.
        for (int i = 0; i < 3; i++) ok = ok && U[i][i] != 0.0;
        double Ui[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (ok) {
            for (int c = 2; c >= 0; c--)
                for (int col = 0; col < 3; col++) {
                    double s = (c == col) ? 1.0 : 0.0;
                    for (int k = c + 1; k < 3; k++) s -= U[c][k] * Ui[k][col];
                    Ui[c][col] = s / U[c][c];
                }
        }
        for (int c = 0; c < 3; c++) {
            double s = 0.0;
            for (int i = 0; i < 3; i++) s += Ui[i][c] * Ui[i][c];
            hw[c] = sqrt(FT_CUT * s);
        }
    }
    const float *kv[3] = {Kg.k0, Kg.k1, Kg.k2};
    const int m[3] = {Kg.m0, Kg.m1, Kg.m2};
    for (int c = 0; c < 3; c++) {
        if (!ok) {
            rec.lo[c] = 0;
            rec.hi[c] = m[c] - 1;
        } else {
            const double pad = 1e-5 * hw[c] + 1e-7;
            rec.lo[c] = lower_idx(kv[c], m[c], -hw[c] - pad);
            rec.hi[c] = upper_idx(kv[c], m[c], hw[c] + pad);
        }
    }
    recs[j] = rec;
}

__device__ inline void ckahan(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// phase in turns of x . k for one axis: exact product split, integer part dropped
__device__ inline float turns(float xh, float xl, float k, float &err)
{
    const float p = xh * k;
    const float e = fmaf(xh, k, -p);
    err += fmaf(xl, k, e);
    return p - rintf(p);
}

// Block = 16 waves on a brick of FT_ROWS (k0, one row per wave of a group)
// x 4 (k1, per lane) x 64 (k2, lane) samples.  Every atom's envelope is
// centred at k = 0, so a brick near the origin lists nearly all atoms and one
// at the edge few: the block's FT_GROUPS wave groups take every FT_GROUPS-th
// record of the staged list each, and the groups' compensated sums are added
// in group order at the end (deterministic).  Blocks are numbered bricks
// nearest the grid centre first (Morton order over the centre-out brick
// ranks), and the grid holds several times the blocks that fit at once, so
// the light edge bricks fill the tail.  (2 rows x 8 groups measured best of
// 1 x 16, 2 x 8 and 4 x 4.)
constexpr int FT_ROWS = 2, FT_GROUPS = 16 / FT_ROWS, FT_THREADS = 16 * 64, FCHUNK = 256;

__global__ __launch_bounds__(FT_THREADS) void ft_grid_kernel(const FtRec *__restrict__ recs, int N, gt_kgrid Kg,
                                                             float2 *__restrict__ out, int e0, int e1, int e2)
{
    __shared__ FtRec list[FCHUNK];
    __shared__ int wcnt[FCHUNK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = wave / FT_ROWS, w2 = wave % FT_ROWS;
    const int nb0 = (Kg.m0 + FT_ROWS - 1) / FT_ROWS, nb1 = (Kg.m1 + 3) / 4, nb2 = (Kg.m2 + 63) / 64;
    // Morton index -> centre-out ranks (t0, t1, t2); ranks beyond the brick counts are empty blocks
    const int e[3] = {e0, e1, e2};
    int t[3];
    morton_ranks(blockIdx.x, e, t);
    if (t[0] >= nb0 || t[1] >= nb1 || t[2] >= nb2) return;  // whole block
    const int bb0 = centre_out(t[0], nb0), bb1 = centre_out(t[1], nb1), bb2 = centre_out(t[2], nb2);
    const int i0 = bb0 * FT_ROWS + w2, i1b = bb1 * 4, i2 = bb2 * 64 + lane;
    const int blo[3] = {bb0 * FT_ROWS, i1b, bb2 * 64};
    const int bhi[3] = {min(bb0 * FT_ROWS + FT_ROWS - 1, Kg.m0 - 1), min(i1b + 3, Kg.m1 - 1), min(bb2 * 64 + 63, Kg.m2 - 1)};
    const float k0 = Kg.k0[min(i0, Kg.m0 - 1)];
    const float k2 = Kg.k2[min(i2, Kg.m2 - 1)];
    float k1[4], sr[4], cr[4], si[4], ci[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        k1[q] = Kg.k1[min(i1b + q, Kg.m1 - 1)];
        sr[q] = cr[q] = si[q] = ci[q] = 0.0f;
    }
    const float cut = (float)(FT_CUT * HALF_LOG2E_SQRT * HALF_LOG2E_SQRT);
    for (int base = 0; base < N; base += FCHUNK) {
        // stage the chunk's records that reach the brick (waves 0..3 scan, atom order kept)
        const int j = base + (int)threadIdx.x;
        bool hit = false;
        if (threadIdx.x < FCHUNK && j < N) {
            const FtRec &r = recs[j];
            hit = r.lo[0] <= bhi[0] && r.hi[0] >= blo[0] && r.lo[1] <= bhi[1] && r.hi[1] >= blo[1] &&
                  r.lo[2] <= bhi[2] && r.hi[2] >= blo[2];
        }
        const unsigned long long msk = __ballot(hit);
        if (wave < FCHUNK / 64 && lane == 0) wcnt[wave] = __popcll(msk);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < FCHUNK / 64; w++) {
            if (w < wave) off += wcnt[w];
            total += wcnt[w];
        }
        if (hit) {
            const int4 *src = reinterpret_cast<const int4 *>(recs + j);
            int4 *dst = reinterpret_cast<int4 *>(list + off + __popcll(msk & ((1ull << lane) - 1ull)));
#pragma unroll
            for (int k = 0; k < 8; k++) dst[k] = src[k];
        }
        __syncthreads();
        for (int e = grp; e < total; e += FT_GROUPS) {
            const FtRec &r = list[e];
            if (r.lo[0] > i0 || r.hi[0] < i0) continue;
            float err0 = 0.0f;
            const float t02 = turns(r.xh[0], r.xl[0], k0, err0) + turns(r.xh[2], r.xl[2], k2, err0);
            if (r.iso != 0.0f) {
                const float base2 = fmaf(k0, k0, k2 * k2);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float nq = fmaf(k1[q], k1[q], base2) * (r.u[0] * r.u[0]);
                    if (nq < cut) {
                        float err = err0;
                        const float tt = t02 + turns(r.xh[1], r.xl[1], k1[q], err) + err;
                        float sn, cs;
                        ft_sincos_turns(tt, sn, cs);
                        const float a = r.amp * __builtin_amdgcn_exp2f(-nq);
                        ckahan(sr[q], cr[q], a * cs);
                        ckahan(si[q], ci[q], -a * sn);
                    }
                }
            } else {
                // (U^T k)_0 = u00 k0, (U^T k)_1 = u01 k0 + u11 k1, (U^T k)_2 = u02 k0 + u12 k1 + u22 k2
                const float v0 = r.u[0] * k0;
                const float a1 = r.u[3] * k0;
                const float a2 = fmaf(r.u[5], k0, r.u[2] * k2);
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const float v1 = fmaf(r.u[1], k1[q], a1);
                    const float v2 = fmaf(r.u[4], k1[q], a2);
                    const float nq = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
                    if (nq < cut) {
                        float err = err0;
                        const float tt = t02 + turns(r.xh[1], r.xl[1], k1[q], err) + err;
                        float sn, cs;
                        ft_sincos_turns(tt, sn, cs);
                        const float a = r.amp * __builtin_amdgcn_exp2f(-nq);
                        ckahan(sr[q], cr[q], a * cs);
                        ckahan(si[q], ci[q], -a * sn);
                    }
                }
            }
        }
        __syncthreads();
    }
    // groups' sums (value minus compensation) through LDS, added in group order
    float2 *red = reinterpret_cast<float2 *>(list);  // [group][w2][q][lane]: 8 x 2 x 4 x 64 float2 = 32 KB
    static_assert(sizeof(list) >= FT_GROUPS * FT_ROWS * 4 * 64 * sizeof(float2), "reduction buffer");
#pragma unroll
    for (int q = 0; q < 4; q++) red[((grp * FT_ROWS + w2) * 4 + q) * 64 + lane] = make_float2(sr[q] - cr[q], si[q] - ci[q]);
    __syncthreads();
    if (grp == 0 && i0 < Kg.m0 && i2 < Kg.m2)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (i1b + q >= Kg.m1) continue;
            float tr = 0.0f, ctr = 0.0f, ti = 0.0f, cti = 0.0f;
            for (int g = 0; g < FT_GROUPS; g++) {
                const float2 v = red[((g * FT_ROWS + w2) * 4 + q) * 64 + lane];
                ckahan(tr, ctr, v.x);
                ckahan(ti, cti, v.y);
            }
            out[((size_t)i0 * Kg.m1 + i1b + q) * Kg.m2 + i2] = make_float2(tr, ti);
        }
}

__global__ void cfill_kernel(float2 *p, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = make_float2(0.0f, 0.0f);
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_ft_workspace(const gt_atoms *atoms, const gt_kgrid *grid)
{
    (void)grid;
    if (!atoms) return 0;
    return align_up((size_t)atoms->n * sizeof(FtRec));
}

extern "C" int gt_ft_eval(const gt_atoms *atoms, const gt_kgrid *grid, float *out, void *ws, size_t ws_bytes,
                          void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms))) return rc;
    GT_CHECK_ARG(atoms->dim == 3, "gt_ft_eval: only 3D atoms are supported, got dim %d", atoms->dim);
    GT_CHECK_ARG(grid != nullptr && grid->dim == 3, "gt_ft_eval: grid must be a 3D separable grid");
    GT_CHECK_ARG(grid->m0 >= 1 && grid->m1 >= 1 && grid->m2 >= 1, "gt_ft_eval: grid sizes must be >= 1");
    GT_CHECK_ARG(grid->m0 < (1 << 24) && grid->m1 < (1 << 24) && grid->m2 < (1 << 24), "gt_ft_eval: grid too large");
    GT_CHECK_ARG(grid->k0 && grid->k1 && grid->k2, "gt_ft_eval: grid vectors must be non-NULL");
    GT_CHECK_ARG(out != nullptr, "gt_ft_eval: out is NULL");
    const size_t need = gt_ft_workspace(atoms, grid);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_ft_eval: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 127) == 0, "gt_ft_eval: workspace must be 128-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const size_t nout = (size_t)grid->m0 * grid->m1 * grid->m2;
    if (atoms->n == 0) {
        hipLaunchKernelGGL(cfill_kernel, dim3((unsigned)std::min<size_t>((nout + 255) / 256, 65535)), dim3(256), 0,
                           st, (float2 *)out, nout);
        GT_HIP(hipGetLastError());
        return 0;
    }
    FtRec *recs = (FtRec *)ws;
    hipLaunchKernelGGL(ft_setup_kernel, dim3((atoms->n + 255) / 256), dim3(256), 0, st, *atoms, *grid, recs);
    // bricks of 2 x 4 x 64 samples, numbered by Morton index over 2^e0 x 2^e1 x 2^e2 ranks
    const int e[3] = {ceil_log2((grid->m0 + FT_ROWS - 1) / FT_ROWS), ceil_log2((grid->m1 + 3) / 4),
                      ceil_log2((grid->m2 + 63) / 64)};
    GT_CHECK_ARG(e[0] + e[1] + e[2] <= 30, "gt_ft_eval: grid too large");
    const unsigned blocks = 1u << (e[0] + e[1] + e[2]);
    prof_begin(st, GT_K_FT);
    hipLaunchKernelGGL(ft_grid_kernel, dim3(blocks), dim3(FT_THREADS), 0, st, recs, atoms->n, *grid, (float2 *)out,
                       e[0], e[1], e[2]);
    prof_end(st, GT_K_FT);
    GT_HIP(hipGetLastError());
    return 0;
}
