// gt_kl.hip -- per-angle moments of a sinogram-shaped function f for the KL
// moment initialiser (KL_GaussRadon.py:57-134, 137-190).
//
// The initialiser fits one atom to a residual sinogram by matching, over all
// angles, the mass, the centre and the spread of f against the atom's
// projection.  Everything it needs from the sinogram is, per angle a,
//   m[a] = ( sum f, sum f p0, sum f p1, sum f p0^2, sum f p1^2 )
// over the detector pixels (MOM0, MOM1 and MOM2 of KL_GaussRadon.py:221-392
// are linear combinations of these; 2D: the p1 terms are 0).  f is not stored:
// it is formed per pixel from the residual S and the atom's projection Q,
//   GT_KL_CLAMP     f = max(0, a S + b Q)                 (the ProjGD step, :128)
//   GT_KL_LAGRANGE  f + a log f = S + a log(b Q), f >= 0  (Newton_sino_2D, :193-229;
//                   f = S when a == 0, f = 0 where b Q <= 1e-15)
// so one pass reads S (and Q) once.  FP64 accumulation; block partials over
// 4096-pixel chunks are summed in a fixed order by a second kernel, so the
// result does not depend on scheduling.
#include "gt_common.h"

namespace gt {

constexpr int KM_THREADS = 256;
constexpr int KM_CHUNK = 4096;
constexpr int KM_NM = 5;

// Per-pixel solve of x + t log x = c0 with c0 = s + t log q (t > 0, q > 0):
// Newton on x directly when c0 > 0; otherwise x = t X with X e^X = e^(c0/t)/t
// (Lambert W), Newton from the root of X + X^2 = e^(c0/t)/t.  Ten steps each.
__device__ inline float km_lagrange(float s, float q, float t)
{
    if (t == 0.0f) return s;
    if (!(q > 1e-15f)) return 0.0f;
    float c = fmaf(t, logf(q), s);
    if (c > 0.0f) {
        c += t;
        float X = c / (1.0f + t);
#pragma unroll 1
        for (int it = 0; it < 10; it++) X = X * (c - t * logf(X)) / (X + t);
        return X;
    }
    const float w = expf(c / t) / t;
    float X = 0.5f * (sqrtf(fmaf(4.0f, w, 1.0f)) - 1.0f);
#pragma unroll 1
    for (int it = 0; it < 10; it++) X = fmaf(X, X, w * expf(-X)) / (1.0f + X);
    return X * t;
}

template <int MODE, bool VEC>
__global__ __launch_bounds__(KM_THREADS) void km_partial_kernel(const float *__restrict__ S,
                                                                const float *__restrict__ Q, float a, float b,
                                                                const float *__restrict__ p0,
                                                                const float *__restrict__ p1, int n0, int n1,
                                                                int chunks, double *__restrict__ part)
{
    const int ang = blockIdx.y, c = blockIdx.x, tid = threadIdx.x;
    const size_t npx = (size_t)n0 * n1;
    const float *Sa = S + (size_t)ang * npx;
    const float *Qa = Q ? Q + (size_t)ang * npx : nullptr;
    const int beg = c * KM_CHUNK;
    const int end = (int)min(npx, (size_t)beg + KM_CHUNK);
    double m[KM_NM] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (VEC) {
        // rows of a multiple of 4 pixels, 16-byte aligned: 4 pixels of one row per load
        const float4 *S4 = (const float4 *)Sa;
        const float4 *Q4 = (const float4 *)Qa;
        for (int i = beg + 4 * tid; i < end; i += 4 * KM_THREADS) {
            const int j = i / n1, k = i - j * n1;
            const float4 s = S4[i >> 2];
            const float4 q = Qa ? Q4[i >> 2] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float4 y1 = p1 ? *(const float4 *)(p1 + k) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const float sv[4] = {s.x, s.y, s.z, s.w}, qv[4] = {q.x, q.y, q.z, q.w};
            const float yv[4] = {y1.x, y1.y, y1.z, y1.w};
            double row = 0.0, r1 = 0.0, r11 = 0.0;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float f = (MODE == GT_KL_CLAMP) ? fmaxf(0.0f, fmaf(a, sv[u], b * qv[u]))
                                                      : km_lagrange(sv[u], b * qv[u], a);
                const double fd = f, y = yv[u];
                row += fd;
                r1 = fma(fd, y, r1);
                r11 = fma(fd * y, y, r11);
            }
            const double y0 = p0[j];
            m[0] += row;
            m[1] = fma(row, y0, m[1]);
            m[2] += r1;
            m[3] = fma(row * y0, y0, m[3]);
            m[4] += r11;
        }
    } else {
        for (int i = beg + tid; i < end; i += KM_THREADS) {
            const int j = i / n1, k = i - j * n1;
            const float s = Sa[i];
            const float q = Qa ? Qa[i] : 0.0f;
            float f;
            if (MODE == GT_KL_CLAMP)
                f = fmaxf(0.0f, fmaf(a, s, b * q));
            else
                f = km_lagrange(s, b * q, a);
            const double fd = f, y0 = p0[j], y1 = p1 ? (double)p1[k] : 0.0;
            m[0] += fd;
            m[1] = fma(fd, y0, m[1]);
            m[2] = fma(fd, y1, m[2]);
            m[3] = fma(fd * y0, y0, m[3]);
            m[4] = fma(fd * y1, y1, m[4]);
        }
    }
    __shared__ double red[KM_THREADS / 64][KM_NM];
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int q = 0; q < KM_NM; q++) {
        double v = m[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
        if (lane == 0) red[wave][q] = v;
    }
    __syncthreads();
    if (tid < KM_NM) {
        double v = 0.0;
        for (int w = 0; w < KM_THREADS / 64; w++) v += red[w][tid];
        part[((size_t)ang * chunks + c) * KM_NM + tid] = v;
    }
}

__global__ void km_finish_kernel(const double *__restrict__ part, int chunks, int nA, double *__restrict__ out)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nA * KM_NM) return;
    const int ang = idx / KM_NM, q = idx - ang * KM_NM;
    double s = 0.0;
    for (int c = 0; c < chunks; c++) s += part[((size_t)ang * chunks + c) * KM_NM + q];
    out[idx] = s;
}

static int km_chunks(const gt_geom *g) { return (int)(((size_t)g->n0 * g->n1 + KM_CHUNK - 1) / KM_CHUNK); }

}  // namespace gt

using namespace gt;

extern "C" size_t gt_sino_moments_workspace(const gt_geom *geom)
{
    if (!geom) return 0;
    return align_up((size_t)geom->nA * km_chunks(geom) * KM_NM * sizeof(double));
}

extern "C" int gt_sino_moments(const gt_geom *geom, const float *S, const float *Q, int mode, double a, double b,
                               double *out, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    GT_CHECK_ARG(geom != nullptr, "geometry is NULL");
    if ((rc = check_geom(geom, geom->dim))) return rc;
    GT_CHECK_ARG(S && out, "S and out must be non-NULL");
    GT_CHECK_ARG(mode == GT_KL_CLAMP || mode == GT_KL_LAGRANGE, "mode must be GT_KL_CLAMP or GT_KL_LAGRANGE");
    GT_CHECK_ARG(mode == GT_KL_CLAMP || Q != nullptr, "GT_KL_LAGRANGE needs Q");
    GT_CHECK_ARG(mode == GT_KL_CLAMP || a >= 0.0, "GT_KL_LAGRANGE needs a >= 0");
    GT_CHECK_ARG((size_t)geom->n0 * geom->n1 < ((size_t)1 << 31), "detector too large: %d x %d", geom->n0,
                 geom->n1);
    const size_t need = gt_sino_moments_workspace(geom);
    GT_CHECK_ARG(ws != nullptr || need == 0, "workspace is NULL");
    if (ws_bytes < need) {
        set_error("workspace too small: %zu < %zu", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    hipStream_t st = (hipStream_t)stream;
    const int chunks = km_chunks(geom);
    double *part = (double *)ws;
    const float *p1 = geom->dim == 3 ? geom->p1 : nullptr;
    dim3 grid(chunks, geom->nA);
    // 16-byte loads when every row is a multiple of 4 pixels and the arrays are 16-byte aligned
    auto al16 = [](const void *p) { return ((uintptr_t)p & 15) == 0; };
    const bool vec = geom->dim == 3 && geom->n1 % 4 == 0 && al16(S) && (Q == nullptr || al16(Q)) && al16(p1);
#define GT_KM_LAUNCH(M, V)                                                                                     \
    hipLaunchKernelGGL((km_partial_kernel<M, V>), grid, dim3(KM_THREADS), 0, st, S, Q, (float)a, (float)b,   \
                       geom->p0, p1, geom->n0, geom->n1, chunks, part)
    if (mode == GT_KL_CLAMP) {
        if (vec)
            GT_KM_LAUNCH(GT_KL_CLAMP, true);
        else
            GT_KM_LAUNCH(GT_KL_CLAMP, false);
    } else {
        if (vec)
            GT_KM_LAUNCH(GT_KL_LAGRANGE, true);
        else
            GT_KM_LAUNCH(GT_KL_LAGRANGE, false);
    }
#undef GT_KM_LAUNCH
    GT_HIP(hipGetLastError());
    const int n = geom->nA * KM_NM;
    hipLaunchKernelGGL(km_finish_kernel, dim3((n + 255) / 256), dim3(256), 0, st, part, chunks, geom->nA, out);
    GT_HIP(hipGetLastError());
    return 0;
}
