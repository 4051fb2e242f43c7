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
// Per-point complex sums are taken in FP64.
#include "gt_common.h"
#include "gt_ftpatch.h"

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
    // radius of a ball about k = 0 holding the support's box (+inf: no bound)
    rec.pad0[0] = ok ? (float)(sqrt(hw[0] * hw[0] + hw[1] * hw[1] + hw[2] * hw[2]) * (1.0 + 1e-5) + 1e-6)
                     : __builtin_huge_valf();
    recs[j] = rec;
}

// phase in turns of x . k for one axis: exact product split, integer part dropped
__device__ inline float turns(float xh, float xl, float k, float &err)
{
    const float p = xh * k;
    const float e = fmaf(xh, k, -p);
    err += fmaf(xl, k, e);
    return p - rintf(p);
}

// ---------------------------------------------------------------------------
// Every atom's support {|U^T k|^2 < 30} is an ellipsoid about k = 0 (about 1%
// of a 128^3 grid), so the samples near the origin see nearly every atom and
// the outer ones few.  The records are first put in order of decreasing
// radius of a ball about k = 0 holding the support (ft_rank_kernel, ties by
// atom index); a patch whose nearest sample lies at distance d from the
// origin then only looks at the leading records of radius >= d.
constexpr int FT_SORT_MAX = 16384;  // atoms ranked by radius; more are taken in atom order

// One wave per atom j: its rank = the atoms of larger radius, or of equal
// radius and smaller index; the record is copied to that position.
__global__ __launch_bounds__(256) void ft_rank_kernel(const FtRec *__restrict__ recs, int N, FtRec *__restrict__ sorted)
{
    const int lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= N) return;
    const float rj = recs[j].pad0[0];
    int rank = 0;
    for (int i = lane; i < N; i += 64) {
        const float ri = recs[i].pad0[0];
        rank += (ri > rj || (ri == rj && i < j)) ? 1 : 0;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) rank += __shfl_xor(rank, o);
    // eight lanes copy the 128-byte record, 16 bytes each
    if (lane < 8)
        reinterpret_cast<int4 *>(sorted + rank)[lane] = reinterpret_cast<const int4 *>(recs + j)[lane];
}

// Patch kernel: one block of 4 waves per 8 x 8 (k1, k2) patch at one k0;
// lane l holds the sample at offset (l / 8, l % 8).  Staging (a thread per
// record, records of radius >= the patch's distance from the origin only):
// the record's index box against the patch and, for an ascending grid, the
// least |U^T k|^2 over the patch against the cutoff (ft_patch_min, with a
// relative slack of 1e-4 so that no sample the FP32 test takes is lost);
// records that meet the patch go to LDS in record order.  Wave w evaluates
// the staged records w, w + 4, ... (branch-free: a sample outside the support
// adds an exact zero) summing in FP64, and the 4 partial sums are added in
// wave order: every sample's sum has an order fixed by the record order.
constexpr int FP_THREADS = 256, FP_WAVES = 4, FP_CHUNK = 256, FP_PATCH = 8;

__global__ __launch_bounds__(FP_THREADS) void ft_patch_kernel(const FtRec *__restrict__ recs, int N, gt_kgrid Kg,
                                                              float2 *__restrict__ out, int e0, int e1, int e2,
                                                              int by_radius)
{
    __shared__ FtL list[FP_CHUNK];  // also the reduction buffer at the end
    __shared__ int wcnt[FP_WAVES];
    static_assert(sizeof(list) >= (size_t)FP_WAVES * 64 * sizeof(double2), "reduction buffer");
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int nb1 = (Kg.m1 + FP_PATCH - 1) / FP_PATCH, nb2 = (Kg.m2 + FP_PATCH - 1) / FP_PATCH;
    const int e[3] = {e0, e1, e2};
    int t[3];
    morton_ranks(blockIdx.x, e, t);
    if (t[0] >= Kg.m0 || t[1] >= nb1 || t[2] >= nb2) return;  // whole block: no patch of this rank
    const int i0 = centre_out(t[0], Kg.m0);
    const int a1 = centre_out(t[1], nb1) * FP_PATCH, a2 = centre_out(t[2], nb2) * FP_PATCH;
    const int z1 = min(a1 + FP_PATCH - 1, Kg.m1 - 1), z2 = min(a2 + FP_PATCH - 1, Kg.m2 - 1);
    const int i1 = a1 + (lane >> 3), i2 = a2 + (lane & 7);
    const float k0 = Kg.k0[i0], k1 = Kg.k1[min(i1, Kg.m1 - 1)], k2 = Kg.k2[min(i2, Kg.m2 - 1)];
    const float c1a = Kg.k1[a1], c1b = Kg.k1[z1], c2a = Kg.k2[a2], c2b = Kg.k2[z2];
    const bool exact = Kg.sorted != 0;
    // distance from k = 0 to the patch (its samples lie in the coordinate box)
    const float g1 = (c1a <= 0.0f && c1b >= 0.0f) ? 0.0f : fminf(fabsf(c1a), fabsf(c1b));
    const float g2 = (c2a <= 0.0f && c2b >= 0.0f) ? 0.0f : fminf(fabsf(c2a), fabsf(c2b));
    const float dmin = exact ? sqrtf(fmaf(k0, k0, fmaf(g1, g1, g2 * g2))) : 0.0f;
    const float cut = (float)(FT_CUT * HALF_LOG2E_SQRT * HALF_LOG2E_SQRT);
    const float cut_keep = cut * (1.0f + 1e-4f);
    double sr = 0.0, si = 0.0;
    for (int base = 0; base < N; base += FP_CHUNK) {
        // records in radius order: this and every later one fall short of the patch
        if (by_radius && recs[base].pad0[0] < dmin) break;
        const int j = base + (int)threadIdx.x;
        bool hit = false;
        if (j < N) {
            const FtL &rec = *reinterpret_cast<const FtL *>(recs + j);
            const int *bx = reinterpret_cast<const int *>(recs + j) + 16;  // lo[3], hi[3]
            hit = rec.pad[0] >= dmin && bx[0] <= i0 && bx[3] >= i0 && bx[1] <= z1 && bx[4] >= a1 && bx[2] <= z2 &&
                  bx[5] >= a2 && (!exact || ft_patch_min(rec, k0, c1a, c1b, c2a, c2b) < cut_keep);
        }
        const unsigned long long m = __ballot(hit);
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int off = 0, total = 0;
#pragma unroll
        for (int w = 0; w < FP_WAVES; w++) {
            if (w < wave) off += wcnt[w];
            total += wcnt[w];
        }
        if (hit) list[off + __popcll(m & ((1ull << lane) - 1ull))] = *reinterpret_cast<const FtL *>(recs + j);
        __syncthreads();
        for (int q = wave; q < total; q += FP_WAVES) {
            const FtL &r = list[q];
            float nq;
            if (r.iso != 0.0f) {
                nq = fmaf(k1, k1, fmaf(k0, k0, k2 * k2)) * (r.u[0] * r.u[0]);
            } else {
                const float v0 = r.u[0] * k0;
                const float v1 = fmaf(r.u[1], k1, r.u[3] * k0);
                const float v2 = fmaf(r.u[4], k1, fmaf(r.u[5], k0, r.u[2] * k2));
                nq = fmaf(v0, v0, fmaf(v1, v1, v2 * v2));
            }
            float err = 0.0f;
            const float t01 = turns(r.xh[0], r.xl[0], k0, err) + turns(r.xh[1], r.xl[1], k1, err);
            const float tt = (t01 + turns(r.xh[2], r.xl[2], k2, err)) + err;
            float sn, cs;
            ft_sincos_turns(tt, sn, cs);
            const float a = nq < cut ? r.amp * __builtin_amdgcn_exp2f(-nq) : 0.0f;
            sr += (double)(a * cs);
            si += (double)(-a * sn);
        }
        __syncthreads();
    }
    double2 *red = reinterpret_cast<double2 *>(list);  // [wave][lane]
    red[wave * 64 + lane] = make_double2(sr, si);
    __syncthreads();
    if (wave == 0 && i1 < Kg.m1 && i2 < Kg.m2) {
        double tr = 0.0, ti = 0.0;
        for (int w = 0; w < FP_WAVES; w++) {
            const double2 v = red[w * 64 + lane];
            tr += v.x;
            ti += v.y;
        }
        out[((size_t)i0 * Kg.m1 + i1) * Kg.m2 + i2] = make_float2((float)tr, (float)ti);
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
    // the records, then (for at most FT_SORT_MAX atoms) their copy in radius order
    const size_t one = align_up((size_t)atoms->n * sizeof(FtRec));
    return atoms->n <= FT_SORT_MAX ? 2 * one : one;
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
    const unsigned sb = (unsigned)((atoms->n + 255) / 256);
    hipLaunchKernelGGL(ft_setup_kernel, dim3(sb), dim3(256), 0, st, *atoms, *grid, recs);
    const FtRec *order = recs;
    if (atoms->n <= FT_SORT_MAX) {
        FtRec *sorted = (FtRec *)((char *)ws + align_up((size_t)atoms->n * sizeof(FtRec)));
        hipLaunchKernelGGL(ft_rank_kernel, dim3((unsigned)((atoms->n + 3) / 4)), dim3(256), 0, st, recs, atoms->n,
                           sorted);
        order = sorted;
    }
    // patches of 1 x 8 x 8 samples, numbered by Morton index over 2^e0 x 2^e1 x 2^e2 centre-out ranks
    const int e[3] = {ceil_log2(grid->m0), ceil_log2((grid->m1 + FP_PATCH - 1) / FP_PATCH),
                      ceil_log2((grid->m2 + FP_PATCH - 1) / FP_PATCH)};
    GT_CHECK_ARG(e[0] + e[1] + e[2] <= 30, "gt_ft_eval: grid too large");
    const unsigned blocks = 1u << (e[0] + e[1] + e[2]);
    prof_begin(st, GT_K_FT);
    hipLaunchKernelGGL(ft_patch_kernel, dim3(blocks), dim3(FP_THREADS), 0, st, order, atoms->n, *grid,
                       (float2 *)out, e[0], e[1], e[2], order != recs ? 1 : 0);
    prof_end(st, GT_K_FT);
    GT_HIP(hipGetLastError());
    return 0;
}
