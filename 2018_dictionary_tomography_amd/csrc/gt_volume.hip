// gt_volume.hip -- volume render of Gaussian atoms (GaussTomo.discretise).
//
// u(y) = sum_atoms I sqrt(30/pi) exp(-30 |U (y - x)|^2) [|U (y - x)|^2 <= 1]
// (__f / __VolProj, GaussDictCode/bin/NewtonBases.py:237-257, 698-711 and
// NewtonBases2D.py:16-34, 259-271).
//
// One FP64 setup thread per atom writes a VolRec: amplitude, centre as a
// float pair, U scaled so that the exponent is exp2(-|U'(y-x)|^2), and the
// index box of the ellipsoid |U(y-x)| <= 1 on each (ascending) coordinate
// vector.  The render kernel gives each block a brick of voxels, stages the
// atoms whose box meets the brick in LDS (in atom order), and accumulates
// 4 voxels per lane with Kahan sums.
#include "gt_common.h"

namespace gt {

struct VolRec {
    float amp;
    float xh[3], xl[3];
    float u[6];  // scaled U slots [u00, u11, u22, u01, u12, u02]
    float pad0[3];
    int lo[3], hi[3];
    int pad1[10];
};
static_assert(sizeof(VolRec) == 128, "VolRec must stay 128 bytes");

constexpr double SQRT_NEG_KAPPA = 6.578818071009935;  // sqrt(30 log2 e)

template <int D> __global__ __launch_bounds__(256) void vol_setup_kernel(gt_atoms A, gt_vol V, VolRec *recs)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    constexpr int NR = Dims<D>::NR;
    PairGeom<D> g;
    load_atom<D>(A, j, g);
    VolRec rec;
    rec.amp = (float)(g.I * sqrt(30.0 / 3.14159265358979323846));
    for (int i = 0; i < 3; i++) {
        const double xv = (i < D) ? g.x[i] : 0.0;
        rec.xh[i] = (float)xv;
        rec.xl[i] = (float)(xv - (double)rec.xh[i]);
        rec.pad0[i] = 0.0f;
    }
    for (int k = 0; k < 6; k++) rec.u[k] = 0.0f;
    for (int k = 0; k < NR; k++) rec.u[k] = (float)(SQRT_NEG_KAPPA * g.U[slot_row(D, k)][slot_col(D, k)]);
    for (int k = 0; k < 10; k++) rec.pad1[k] = 0;
    const float *y[3] = {V.y0, V.y1, V.y2};
    const int m[3] = {V.m0, V.m1, V.m2};
    // Half-widths of {|U(y-x)| <= 1}: norms of the rows of U^{-1}, i.e. of the
    // columns of U^{-T}; U is upper triangular.
    double Ui[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool ok = V.sorted != 0;
    for (int i = 0; i < D; i++) ok = ok && g.U[i][i] != 0.0;
    if (ok) {
        for (int c = D - 1; c >= 0; c--) {  // back substitution for U Ui = Id
            for (int col = 0; col < D; col++) {
                double s = (c == col) ? 1.0 : 0.0;
                for (int k = c + 1; k < D; k++) s -= g.U[c][k] * Ui[k][col];
                Ui[c][col] = s / g.U[c][c];
            }
        }
    }
    for (int c = 0; c < 3; c++) {
        if (c >= D) {
            rec.lo[c] = 0;
            rec.hi[c] = 0;
            continue;
        }
        if (!ok) {
            rec.lo[c] = 0;
            rec.hi[c] = m[c] - 1;
            continue;
        }
        double hw = 0.0;
        for (int k = 0; k < D; k++) hw += Ui[c][k] * Ui[c][k];
        hw = sqrt(hw);
        const double pad = 1e-5 * hw + 1e-7;
        rec.lo[c] = lower_idx(y[c], m[c], g.x[c] - hw - pad);
        rec.hi[c] = upper_idx(y[c], m[c], g.x[c] + hw + pad);
    }
    recs[j] = rec;
}

__device__ inline void vkahan(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// Stage the atoms whose box meets [b_lo, b_hi] (3 axes), keeping atom order.
__device__ inline int vol_stage(const VolRec *__restrict__ recs, int base, int N, const int *b_lo, const int *b_hi,
                                VolRec *list, int *wcnt)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = base + (int)threadIdx.x;
    bool hit = false;
    if (j < N) {
        const VolRec &r = recs[j];
        hit = r.lo[0] <= b_hi[0] && r.hi[0] >= b_lo[0] && r.lo[1] <= b_hi[1] && r.hi[1] >= b_lo[1] &&
              r.lo[2] <= b_hi[2] && r.hi[2] >= b_lo[2];
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
    if (hit) {
        const int4 *src = reinterpret_cast<const int4 *>(recs + j);
        int4 *dst = reinterpret_cast<int4 *>(list + off + __popcll(m & ((1ull << lane) - 1ull)));
#pragma unroll
        for (int k = 0; k < 8; k++) dst[k] = src[k];
    }
    __syncthreads();
    return total;
}

constexpr int VCHUNK = 128;

// 3D: block brick = 4 (j0, one per wave) x 4 (j1, per lane) x 64 (j2, lane).
__global__ __launch_bounds__(256) void volume3_kernel(const VolRec *__restrict__ recs, int N, gt_vol V,
                                                      float *__restrict__ u)
{
    __shared__ VolRec list[VCHUNK];
    __shared__ int wcnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb2 = (V.m2 + 63) / 64, nb1 = (V.m1 + 3) / 4;
    int b = blockIdx.x;
    const int bb2 = b % nb2;
    b /= nb2;
    const int bb1 = b % nb1;
    const int bb0 = b / nb1;
    const int j0 = bb0 * 4 + wave, j1b = bb1 * 4, j2 = bb2 * 64 + lane;
    const int blo[3] = {bb0 * 4, j1b, bb2 * 64};
    const int bhi[3] = {min(bb0 * 4 + 3, V.m0 - 1), min(j1b + 3, V.m1 - 1), min(bb2 * 64 + 63, V.m2 - 1)};
    const float y0 = V.y0[min(j0, V.m0 - 1)];
    const float y2 = V.y2[min(j2, V.m2 - 1)];
    float y1[4], s[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        y1[k] = V.y1[min(j1b + k, V.m1 - 1)];
        s[k] = 0.0f;
        c[k] = 0.0f;
    }
    for (int base = 0; base < N; base += VCHUNK) {
        // stage with the first VCHUNK threads only
        const int total = vol_stage(recs, base, min(N, base + VCHUNK), blo, bhi, list, wcnt);
        for (int e = 0; e < total; e++) {
            const VolRec &r = list[e];
            if (r.lo[0] > j0 || r.hi[0] < j0) continue;
            const float d0 = (y0 - r.xh[0]) - r.xl[0];
            const float d2 = (y2 - r.xh[2]) - r.xl[2];
            const float a0 = fmaf(r.u[0], d0, r.u[5] * d2);
            const float a1 = r.u[4] * d2;
            const float u2 = r.u[2] * d2;
            const float c2 = u2 * u2;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float d1 = (y1[k] - r.xh[1]) - r.xl[1];
                const float u0 = fmaf(r.u[3], d1, a0);
                const float u1 = fmaf(r.u[1], d1, a1);
                const float q = -fmaf(u0, u0, fmaf(u1, u1, c2));
                if (q >= (float)KAPPA) vkahan(s[k], c[k], r.amp * __builtin_amdgcn_exp2f(q));
            }
        }
        __syncthreads();
    }
    if (j0 < V.m0 && j2 < V.m2)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (j1b + k < V.m1) u[((size_t)j0 * V.m1 + j1b + k) * V.m2 + j2] = s[k];
}

// 2D: block tile = 16 rows (j0, 4 per lane, 4 waves) x 64 cols (j1, lane).
__global__ __launch_bounds__(256) void volume2_kernel(const VolRec *__restrict__ recs, int N, gt_vol V,
                                                      float *__restrict__ u)
{
    __shared__ VolRec list[VCHUNK];
    __shared__ int wcnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nb1 = (V.m1 + 63) / 64;
    const int bb0 = blockIdx.x / nb1, bb1 = blockIdx.x - bb0 * nb1;
    const int r0 = bb0 * 16 + wave * 4, j1 = bb1 * 64 + lane;
    const int blo[3] = {bb0 * 16, bb1 * 64, 0};
    const int bhi[3] = {min(bb0 * 16 + 15, V.m0 - 1), min(bb1 * 64 + 63, V.m1 - 1), 0};
    const float y1 = V.y1[min(j1, V.m1 - 1)];
    float y0[4], s[4], c[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        y0[k] = V.y0[min(r0 + k, V.m0 - 1)];
        s[k] = 0.0f;
        c[k] = 0.0f;
    }
    for (int base = 0; base < N; base += VCHUNK) {
        const int total = vol_stage(recs, base, min(N, base + VCHUNK), blo, bhi, list, wcnt);
        for (int e = 0; e < total; e++) {
            const VolRec &r = list[e];
            if (r.lo[0] > r0 + 3 || r.hi[0] < r0) continue;
            const float d1 = (y1 - r.xh[1]) - r.xl[1];
            const float u1 = r.u[1] * d1;
            const float c1 = u1 * u1;
            const float a0 = r.u[2] * d1;  // slot 2 of a 2D atom is u01
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float d0 = (y0[k] - r.xh[0]) - r.xl[0];
                const float u0 = fmaf(r.u[0], d0, a0);
                const float q = -fmaf(u0, u0, c1);
                if (q >= (float)KAPPA) vkahan(s[k], c[k], r.amp * __builtin_amdgcn_exp2f(q));
            }
        }
        __syncthreads();
    }
    if (j1 < V.m1)
#pragma unroll
        for (int k = 0; k < 4; k++)
            if (r0 + k < V.m0) u[(size_t)(r0 + k) * V.m1 + j1] = s[k];
}

int launch_fill(float *p, size_t n, float v, hipStream_t st);

static int check_vol(const gt_vol *v, int dim)
{
    GT_CHECK_ARG(v != nullptr, "volume grid is NULL");
    GT_CHECK_ARG(v->dim == dim, "volume.dim (%d) does not match atoms.dim (%d)", v->dim, dim);
    GT_CHECK_ARG(v->m0 >= 1 && v->m1 >= 1 && v->m2 >= 1, "volume sizes must be >= 1");
    GT_CHECK_ARG(v->m0 < (1 << 24) && v->m1 < (1 << 24) && v->m2 < (1 << 24), "volume sizes out of range");
    if (dim == 2) GT_CHECK_ARG(v->m2 == 1, "2D volume needs m2 == 1");
    GT_CHECK_ARG(v->y0 && v->y1 && (dim == 2 || v->y2), "volume coordinate vectors must be non-NULL");
    GT_CHECK_ARG(v->sorted == 0 || v->sorted == 1, "volume.sorted must be 0 or 1");
    return 0;
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_volume_workspace(const gt_atoms *atoms, const gt_vol *vol)
{
    (void)vol;
    if (!atoms) return 0;
    return align_up((size_t)atoms->n * sizeof(VolRec));
}

extern "C" int gt_volume(const gt_atoms *atoms, const gt_vol *vol, float *u, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_vol(vol, atoms->dim))) return rc;
    GT_CHECK_ARG(u != nullptr, "gt_volume: u is NULL");
    const size_t need = gt_volume_workspace(atoms, vol);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_volume: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 127) == 0, "gt_volume: workspace must be 128-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const size_t nout = (size_t)vol->m0 * vol->m1 * vol->m2;
    if (atoms->n == 0) return launch_fill(u, nout, 0.0f, st);
    gt_vol V = *vol;
    if (atoms->dim == 2) V.y2 = V.y1;  // unused third axis
    VolRec *recs = (VolRec *)ws;
    const unsigned sb = (unsigned)((atoms->n + 255) / 256);
    if (atoms->dim == 3) {
        hipLaunchKernelGGL(vol_setup_kernel<3>, dim3(sb), dim3(256), 0, st, *atoms, V, recs);
        const long blocks = (long)((V.m0 + 3) / 4) * ((V.m1 + 3) / 4) * ((V.m2 + 63) / 64);
        GT_CHECK_ARG(blocks < (1L << 31), "gt_volume: grid too large");
        prof_begin(st, GT_K_VOLUME);
        hipLaunchKernelGGL(volume3_kernel, dim3((unsigned)blocks), dim3(256), 0, st, recs, atoms->n, V, u);
        prof_end(st, GT_K_VOLUME);
    } else {
        hipLaunchKernelGGL(vol_setup_kernel<2>, dim3(sb), dim3(256), 0, st, *atoms, V, recs);
        const long blocks = (long)((V.m0 + 15) / 16) * ((V.m1 + 63) / 64);
        prof_begin(st, GT_K_VOLUME);
        hipLaunchKernelGGL(volume2_kernel, dim3((unsigned)blocks), dim3(256), 0, st, recs, atoms->n, V, u);
        prof_end(st, GT_K_VOLUME);
    }
    GT_HIP(hipGetLastError());
    return 0;
}
