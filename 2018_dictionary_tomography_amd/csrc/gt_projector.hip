// gt_projector.hip -- forward projection and per-atom derivatives on CDNA4.
//
// Pipeline (all stream-ordered, no host synchronisation):
//   setup    one thread per (atom, angle): FP64 detector Gaussian -> PairRec
//   forward  one block per (angle, 8x64 pixel tile), centre tiles first:
//            stages the angle's records that touch the tile in LDS; each of
//            4 waves owns an 8x16 rectangle (2 pixels per lane) and walks
//            the records whose support meets it; FP64 pixel sums of FP32
//            values taken four records at a time
//   moments  one wave, and one block, per (atom, angle): FP32 separable
//            pixel moments over the rows of each column chunk that the
//            support can reach, LDS-transposed wave reduction, no atomics
//   finish   one wave per atom, one lane per angle: FP64 moment -> (f, g, H)
//            contraction (gt_pair.h), FP64 wave reduction over angles
//
// Reference counterparts: __RadProj_3D/_2p5D (NewtonBases.py:729-756),
// __RadProj (NewtonBases2D.py:284-296), __L2derivs_RadProj_aniso_3D and
// __derivs_RadProj_aniso_3D (NewtonBases.py:792-870, 993-1069) and their 2D
// versions (NewtonBases2D.py:329-401, 444-515).
#include "gt_pixel.h"

namespace gt {

constexpr int TILE_ROWS = 8, TILE_COLS = 64;
constexpr int CHUNK = 256;
constexpr int MOM_STRIDE = 32;  // floats per (atom, angle) moment record
constexpr int MOM_TROWS = 512;  // rows of a moments wave's row table
constexpr int MOM_WAVES = 1;    // waves (pairs) per moments block
static_assert(MOM_TROWS * 16 <= 64 * (MOM_STRIDE + 1) * 4, "row table must fit the reduction slab");
constexpr int MODE_LINEAR = GT_MODE_LINEAR, MODE_L2 = GT_MODE_L2, MODE_RESID = GT_MODE_RESID;

__device__ inline void kahan_add(float &s, float &c, float v)
{
    const float y = v - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
}

// ---------------------------------------------------------------------------
// setup
// ---------------------------------------------------------------------------
template <int D> __global__ __launch_bounds__(256) void setup_kernel(gt_atoms A, gt_geom G, PairRec *recs)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)A.n * G.nA;
    if (idx >= total) return;
    const int a = (int)(idx / A.n), j = (int)(idx - (long)a * A.n);
    PairGeom<D> g;
    load_atom<D>(A, j, g);
    load_angle<D>(G, a, g);
    const float *det[2] = {G.p0, G.p1};
    const int nd[2] = {G.n0, G.n1};
    PairRec rec;
    pair_setup<D>(g, det, nd, G.sorted, rec);
    recs[idx] = rec;
}

// Deterministic block-wide compaction of the records that touch a box
// [r_lo, r_hi] x [c_lo, c_hi]; returns the list length.  The list keeps atom
// order, so the per-pixel summation order does not depend on timing.
template <int NW>
__device__ inline int stage_chunk(const PairRec *__restrict__ ra, int base, int N, int r_lo, int r_hi, int c_lo,
                                  int c_hi, PairRec *list, int *wcnt)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int j = base + (int)threadIdx.x;
    bool hit = false;
    int4 v0 = make_int4(0, 0, 0, 0), v1 = v0, v2 = v0, v3 = v0;
    if (j < N) {
        const int4 *src = reinterpret_cast<const int4 *>(ra + j);
        v0 = src[0];
        v1 = src[1];
        v2 = src[2];
        v3 = src[3];
        // v3 = (lo[0], lo[1], hi[0], hi[1])
        hit = v3.x <= r_hi && v3.z >= r_lo && v3.y <= c_hi && v3.w >= c_lo;
    }
    const unsigned long long m = __ballot(hit);
    if (lane == 0) wcnt[wave] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NW; w++) {
        if (w < wave) off += wcnt[w];
        total += wcnt[w];
    }
    if (hit) {
        int4 *dst = reinterpret_cast<int4 *>(list + off + __popcll(m & ((1ull << lane) - 1ull)));
        dst[0] = v0;
        dst[1] = v1;
        dst[2] = v2;
        dst[3] = v3;
    }
    __syncthreads();
    return total;
}

// ---------------------------------------------------------------------------
// forward, 2D detectors (3D and 2.5D geometries)
// ---------------------------------------------------------------------------
// Block tile TILE_ROWS x TILE_COLS = 8 x 64 pixels, F3_WAVES = 4 waves; wave
// w owns the 8 x 16 rectangle rows 8 (w / 4) .. +7, columns 16 (w % 4) ..
// +15, and lane l column l % 16 of it, rows l / 16 and l / 16 + 4 (2 pixels
// per lane).  Staging keeps the records that touch the tile in LDS and, for
// every wave, the list of those that touch its rectangle (both in atom
// order, so each pixel's summation order is fixed): a wave then walks only
// records that reach its pixels.  (8 x 64 tiles of 4 waves measured 138 us
// against 155 us for 16 x 64 tiles of 8 waves: twice the blocks, so the
// centre-first dispatch balances finer.)
constexpr int F3_WAVES = 4, F3_CHUNK = F3_WAVES * 64;
static_assert(TILE_ROWS == 8 * ((F3_WAVES + 3) / 4) && TILE_COLS == 16 * (F3_WAVES < 4 ? F3_WAVES : 4),
              "the tile is the union of the waves' 8 x 16 rectangles");
// first and last detector coordinate of each row group / column group of the
// waves' rectangles (held in registers: indexed only by unrolled constants)
constexpr int F3_RX = 2 * ((F3_WAVES + 3) / 4), F3_CY = 2 * (F3_WAVES < 4 ? F3_WAVES : 4);

// The tiles at the detector centre carry several times the records of those
// at its edges, and the grid holds more blocks than fit on the device at once:
// blocks are numbered centre tiles first (centre_out / morton_ranks), all
// angles of one tile together (block b: angle b % nA, tile rank b / nA), so
// the heavy tiles start first and the light ones fill the tail.

__device__ inline int stage_rects(const PairRec *__restrict__ ra, int base, int N, int row0, int col0, int thi0,
                                  int thi1, PairRec *list, unsigned short (*wl)[F3_CHUNK], int (*cnt)[F3_WAVES],
                                  int *wlen, bool exact, const float (&rx)[F3_RX], const float (&cy)[F3_CY])
{
    const int lane = threadIdx.x & 63, hw = threadIdx.x >> 6;
    const int j = base + (int)threadIdx.x;
    int4 v0 = make_int4(0, 0, 0, 0), v1 = v0, v2 = v0, v3 = v0;
    bool hit = false;
    unsigned mask = 0;
    if (j < N) {
        const int4 *src = reinterpret_cast<const int4 *>(ra + j);
        v0 = src[0];
        v1 = src[1];
        v2 = src[2];
        v3 = src[3];
        // v3 = (lo[0], lo[1], hi[0], hi[1])
        hit = v3.x <= thi0 && v3.z >= row0 && v3.y <= thi1 && v3.w >= col0;
        if (hit) {
            // f0 = (alpha, c_hi0, c_hi1, c_lo0), f1 = (c_lo1, q0, q1, q2), f2 = (q3, q4, q5, pad)
            const float4 f0 = *reinterpret_cast<const float4 *>(&v0), f1 = *reinterpret_cast<const float4 *>(&v1);
            const float4 f2 = *reinterpret_cast<const float4 *>(&v2);
#pragma unroll
            for (int w = 0; w < F3_WAVES; w++) {
                const int r = row0 + 8 * (w >> 2), c = col0 + 16 * (w & 3);
                if (v3.x <= r + 7 && v3.z >= r && v3.y <= c + 15 && v3.w >= c &&
                    (!exact || support_meets_rect(f1, f2, (rx[2 * (w >> 2)] - f0.y) - f0.w,
                                               (rx[2 * (w >> 2) + 1] - f0.y) - f0.w, (cy[2 * (w & 3)] - f0.z) - f1.x,
                                               (cy[2 * (w & 3) + 1] - f0.z) - f1.x)))
                    mask |= 1u << w;
            }
        }
    }
    // cnt[k][hw]: hits of this hardware wave, k = 0 the tile list, k = 1 + w wave w's list
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long mt = __ballot(hit);
    if (lane == 0) cnt[0][hw] = __popcll(mt);
    unsigned long long mw[F3_WAVES];
#pragma unroll
    for (int w = 0; w < F3_WAVES; w++) {
        mw[w] = __ballot((mask >> w) & 1u);
        if (lane == 0) cnt[1 + w][hw] = __popcll(mw[w]);
    }
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int h = 0; h < F3_WAVES; h++) {
        if (h < hw) off += cnt[0][h];
        total += cnt[0][h];
    }
    const int pos = off + __popcll(mt & below);
    if (hit) {
        int4 *dst = reinterpret_cast<int4 *>(list + pos);
        dst[0] = v0;
        dst[1] = v1;
        dst[2] = v2;
        dst[3] = v3;
    }
#pragma unroll
    for (int w = 0; w < F3_WAVES; w++) {
        int o = 0, t = 0;
#pragma unroll
        for (int h = 0; h < F3_WAVES; h++) {
            if (h < hw) o += cnt[1 + w][h];
            t += cnt[1 + w][h];
        }
        if ((mask >> w) & 1u) wl[w][o + __popcll(mw[w] & below)] = (unsigned short)pos;
        if (threadIdx.x == 0) wlen[w] = t;
    }
    __syncthreads();
    return total;
}

// One record's values at a lane's two pixels.  The record arrives as its
// first 48 bytes (alpha, c_hi, c_lo, q): f0 = (alpha, c_hi0, c_hi1, c_lo0),
// f1 = (c_lo1, q0, q1, q2), f2 = (q3, q4, q5, pad).  No cut-off test: every
// record in a wave's list reaches its rectangle, and the values at pixels of
// the rectangle outside the support are below e^-30 of the record's peak.
__device__ inline void f3_values(const float4 &f0, const float4 &f1, const float4 &f2, float pc, float prA,
                                 float prB, float &vA, float &vB)
{
    const float dp1 = (pc - f0.z) - f1.x;
    const float rowlin = fmaf(f1.z, dp1, f2.x);
    const float col = fmaf(fmaf(f1.w, dp1, f2.y), dp1, f2.z);
    const float dA = (prA - f0.y) - f0.w;
    const float dB = (prB - f0.y) - f0.w;
    const float qA = fmaf(fmaf(f1.y, dA, rowlin), dA, col);
    const float qB = fmaf(fmaf(f1.y, dB, rowlin), dB, col);
    const float eA = __builtin_amdgcn_exp2f(qA), eB = __builtin_amdgcn_exp2f(qB);
    vA = f0.x * eA;
    vB = f0.x * eB;
}

__global__ __launch_bounds__(F3_WAVES * 64) void forward3_kernel(const PairRec *__restrict__ recs, int N, int nA,
                                                                 int n0, int n1, const float *__restrict__ p0,
                                                                 const float *__restrict__ p1, float *__restrict__ R,
                                                                 const float *__restrict__ sub, int e0, int e1,
                                                                 int sorted)
{
    __shared__ PairRec list[F3_CHUNK];
    __shared__ alignas(16) unsigned short wl[F3_WAVES][F3_CHUNK];  // 8-byte reads of 4 indices
    __shared__ int cnt[F3_WAVES + 1][F3_WAVES];
    __shared__ int wlen[F3_WAVES];
    const int tilesR = (n0 + TILE_ROWS - 1) / TILE_ROWS, tilesC = (n1 + TILE_COLS - 1) / TILE_COLS;
    const int a = (int)(blockIdx.x % (unsigned)nA);
    const int e[3] = {e0, e1, 0};
    int t[3];
    morton_ranks(blockIdx.x / (unsigned)nA, e, t);
    if (t[0] >= tilesR || t[1] >= tilesC) return;  // whole block: no tile of this rank
    const int row0 = centre_out(t[0], tilesR) * TILE_ROWS, col0 = centre_out(t[1], tilesC) * TILE_COLS;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wr0 = row0 + 8 * (wave >> 2), wc0 = col0 + 16 * (wave & 3);
    const int i1 = wc0 + (lane & 15);
    const int rA = wr0 + (lane >> 4), rB = rA + 4;
    const int thi0 = min(row0 + TILE_ROWS - 1, n0 - 1), thi1 = min(col0 + TILE_COLS - 1, n1 - 1);
    const float pc = p1[min(i1, n1 - 1)];
    const float prA = p0[min(rA, n0 - 1)], prB = p0[min(rB, n0 - 1)];
    double sA = 0.0, sB = 0.0;
    // detector coordinates bounding each wave rectangle (first and last
    // row / column), for the exact cull of an ascending detector
    float rx[F3_RX], cy[F3_CY];
#pragma unroll
    for (int w = 0; w < (F3_WAVES + 3) / 4; w++) {
        rx[2 * w] = p0[min(row0 + 8 * w, n0 - 1)];
        rx[2 * w + 1] = p0[min(row0 + 8 * w + 7, n0 - 1)];
    }
#pragma unroll
    for (int w = 0; w < (F3_WAVES < 4 ? F3_WAVES : 4); w++) {
        cy[2 * w] = p1[min(col0 + 16 * w, n1 - 1)];
        cy[2 * w + 1] = p1[min(col0 + 16 * w + 15, n1 - 1)];
    }
    const PairRec *ra = recs + (size_t)a * N;
    for (int base = 0; base < N; base += F3_CHUNK) {
        stage_rects(ra, base, N, row0, col0, thi0, thi1, list, wl, cnt, wlen, sorted != 0, rx, cy);
        const int len = wlen[wave];
        const unsigned short *wlw = wl[wave];
        int e = 0;
        // 4 records per step: their list indices (one 8-byte read) and records
        // (three 16-byte reads each) are all read before any is used; the
        // pixel sums still take the records in list order
        for (; e + 4 <= len; e += 4) {
            const uint2 iv = *reinterpret_cast<const uint2 *>(wlw + e);
            const int idx[4] = {(int)(iv.x & 0xffffu), (int)(iv.x >> 16), (int)(iv.y & 0xffffu), (int)(iv.y >> 16)};
            float4 f0[4], f1[4], f2[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const float4 *rp = reinterpret_cast<const float4 *>(list + idx[u]);
                f0[u] = rp[0];
                f1[u] = rp[1];
                f2[u] = rp[2];
            }
            // the four values of a pixel are summed in FP32 in list order and
            // that sum is added to the pixel's FP64 total: an error of at most
            // 3u times the sum of the four magnitudes, for one conversion and
            // one FP64 add per four records
            float vA[4], vB[4];
#pragma unroll
            for (int u = 0; u < 4; u++) f3_values(f0[u], f1[u], f2[u], pc, prA, prB, vA[u], vB[u]);
            sA += (double)(((vA[0] + vA[1]) + vA[2]) + vA[3]);
            sB += (double)(((vB[0] + vB[1]) + vB[2]) + vB[3]);
        }
        for (; e < len; e++) {
            const float4 *rp = reinterpret_cast<const float4 *>(list + wlw[e]);
            float vA, vB;
            f3_values(rp[0], rp[1], rp[2], pc, prA, prB, vA, vB);
            sA += (double)vA;
            sB += (double)vB;
        }
        __syncthreads();
    }
    // R = projection, or projection - sub (the residual) when sub is given
    if (i1 < n1) {
        const size_t oA = ((size_t)a * n0 + rA) * n1 + i1, oB = ((size_t)a * n0 + rB) * n1 + i1;
        if (rA < n0) R[oA] = sub ? (float)sA - sub[oA] : (float)sA;
        if (rB < n0) R[oB] = sub ? (float)sB - sub[oB] : (float)sB;
    }
}

// ---------------------------------------------------------------------------
// forward, 1D detectors (2D geometry)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void forward2_kernel(const PairRec *__restrict__ recs, int N, int n0,
                                                       const float *__restrict__ p0, float *__restrict__ R,
                                                       const float *__restrict__ sub)
{
    __shared__ PairRec list[CHUNK];
    __shared__ int wcnt[4];
    const int a = blockIdx.y;
    const int row0 = blockIdx.x * 256;
    const int i0 = row0 + (int)threadIdx.x;
    const int wave = threadIdx.x >> 6;
    const int w0 = row0 + wave * 64, w1 = w0 + 63;
    const int thi = min(row0 + 255, n0 - 1);
    const float pr = p0[min(i0, n0 - 1)];
    float s = 0.0f, c = 0.0f;
    const PairRec *ra = recs + (size_t)a * N;
    const float kap = (float)KAPPA;
    for (int base = 0; base < N; base += CHUNK) {
        const int total = stage_chunk<4>(ra, base, N, row0, thi, 0, 0, list, wcnt);
        for (int e = 0; e < total; e++) {
            const PairRec &q = list[e];
            if (q.lo[0] > w1 || q.hi[0] < w0) continue;
            const float dp0 = (pr - q.c_hi[0]) - q.c_lo[0];
            const float qq = fmaf(fmaf(q.q[0], dp0, q.q[3]), dp0, q.q[5]);
            if (qq >= kap) kahan_add(s, c, q.alpha * __builtin_amdgcn_exp2f(qq));
        }
        __syncthreads();
    }
    if (i0 < n0) R[(size_t)a * n0 + i0] = sub ? s - sub[(size_t)a * n0 + i0] : s;
}

// ---------------------------------------------------------------------------
// moments
// ---------------------------------------------------------------------------
// Layout of one moment record (32 floats): [0, NM) sum w1 mono, [15, 15+NM)
// sum E^2 mono, [30] sum_fp (alpha E - C)^2, [31] sum_fp C^2 (L2 mode).
template <int K, int MODE>
__global__ __launch_bounds__(64 * MOM_WAVES) void moments_kernel(const PairRec *__restrict__ recs, int N, int nA,
                                                                 int n0, int n1, const float *__restrict__ p0,
                                                                 const float *__restrict__ p1,
                                                                 const float *__restrict__ C, float *__restrict__ mom)
{
    // the reduction slab of a wave doubles as its row table while it walks
    // its pair (MOM_TROWS float4 = 8 KiB of the 8.25 KiB slab)
    __shared__ alignas(16) float red[MOM_WAVES][64][MOM_STRIDE + 1];
    // the wave index is uniform: scalar pair index, record and slab base
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long pair = (long)blockIdx.x * MOM_WAVES + wave;
    const bool valid = pair < (long)N * nA;
    float acc[MOM_STRIDE];
#pragma unroll
    for (int v = 0; v < MOM_STRIDE; v++) acc[v] = 0.0f;
    if (valid) {
        const PairRec rec = recs[pair];
        const int a = (int)(pair / N);
        if (K == 2)
            pair_moments_rows<MODE, 4, MOM_TROWS>(rec, p0, p1, C + (size_t)a * n0 * n1, n1, lane, acc,
                                                  reinterpret_cast<float4 *>(&red[wave][0][0]));
        else
            pair_moments<K, MODE>(rec, p0, p1, C + (size_t)a * n0 * n1, n1, lane, acc);
    }
    // each wave reduces its own slab: no block barrier, so a wave whose pair
    // is done does not wait for the others of its block
#pragma unroll
    for (int v = 0; v < MOM_STRIDE; v++) red[wave][lane][v] = acc[v];
    wave_sync();
    if (valid) {
        const int col = lane & 31, half = lane >> 5;
        float s = 0.0f;
        for (int i = 0; i < 32; i++) s += red[wave][half * 32 + i][col];
        s += __shfl_xor(s, 32);
        if (lane < 32) mom[pair * MOM_STRIDE + lane] = s;
    }
}

// f = 1/2 sum over angles of the per-angle FP64 sums of squares (one thread).
__global__ void half_total_kernel(const double *__restrict__ per, int n, double *__restrict__ f)
{
    double s = 0.0;
    for (int a = 0; a < n; a++) s += per[a];
    f[0] = 0.5 * s;
}

// FP64 sum of C^2 per angle (the L2 value of the pixels outside every footprint).
__global__ __launch_bounds__(256) void sumsq_kernel(const float *__restrict__ C, long per, double *__restrict__ out)
{
    __shared__ double red[256];
    const int a = blockIdx.x;
    double s = 0.0;
    for (long i = threadIdx.x; i < per; i += 256) {
        const double v = C[(size_t)a * per + i];
        s += v * v;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[a] = red[0];
}

// ---------------------------------------------------------------------------
// finish
// ---------------------------------------------------------------------------
template <int D, int MODE>
__global__ __launch_bounds__(64) void finish_kernel(gt_atoms A, gt_geom G, const PairRec *__restrict__ recs,
                                                    const float *__restrict__ mom, const double *__restrict__ sumC2,
                                                    int want_hess, double *__restrict__ fo, double *__restrict__ go,
                                                    double *__restrict__ Ho)
{
    using Dm = Dims<D>;
    constexpr int P = Dm::P, NM = Dm::NM, NH = P * (P + 1) / 2, NV = 1 + P + NH;
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
        const PairRec rec = recs[pi];
        const float *m = mom + pi * MOM_STRIDE;
        if (MODE == MODE_L2) F += 0.5 * (sumC2[a] - (double)m[31] + (double)m[30]);
        const bool empty = rec.lo[0] > rec.hi[0] || (D == 3 && rec.lo[1] > rec.hi[1]);
        if (empty) continue;
        load_angle<D>(G, a, g);
        double S1[NM], S2[NM];
        for (int k = 0; k < NM; k++) {
            S1[k] = m[k];
            S2[k] = m[15 + k];
        }
        if (MODE == MODE_LINEAR) {
            double t[3], s = 0.0;
            matvec_upper<D>(g.U, g.T, t);
            for (int i = 0; i < D; i++) s += t[i] * t[i];
            F += g.I / sqrt(s) * S1[0];
        }
        const double pstar[2] = {(double)rec.c_hi[0] + (double)rec.c_lo[0],
                                 (double)rec.c_hi[1] + (double)rec.c_lo[1]};
        pair_finish_to<D>(g, pstar, S1, S2, MODE != MODE_LINEAR, want_hess, grad, HessUpper<P>{row + 1 + P});
    }
    row[0] = F;
    for (int i = 0; i < P; i++) row[1 + i] = grad[i];
    __syncthreads();
    for (int col = lane; col < NV; col += 64) {
        double s = 0.0;
        for (int i = 0; i < 64; i++) s += red[i][col];
        red[0][col] = s;  // rows 1..63 of this column are no longer read
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
    {
        // U = rho * Id: d r_k / d rho = 1 for the D diagonal slots.
        constexpr int PI = 2 + D;
        double Hf[P][P];
        {
            int k = 1 + P;
            for (int i = 0; i < P; i++)
                for (int l = i; l < P; l++) {
                    Hf[i][l] = v[k];
                    Hf[l][i] = v[k];
                    k++;
                }
        }
        double gi[PI], Hi[PI][PI];
        double E[P][PI];
        for (int i = 0; i < P; i++)
            for (int l = 0; l < PI; l++) E[i][l] = 0.0;
        for (int i = 0; i <= D; i++) E[i][i] = 1.0;
        for (int k = 0; k < D; k++) E[1 + D + k][1 + D] = 1.0;
        for (int l = 0; l < PI; l++) {
            double s = 0.0;
            for (int i = 0; i < P; i++) s += E[i][l] * v[1 + i];
            gi[l] = s;
        }
        for (int l = 0; l < PI; l++) go[(size_t)j * PI + l] = gi[l];
        if (Ho && want_hess) {
            for (int l = 0; l < PI; l++)
                for (int m2 = 0; m2 < PI; m2++) {
                    double s = 0.0;
                    for (int i = 0; i < P; i++)
                        for (int k = 0; k < P; k++) s += E[i][l] * Hf[i][k] * E[k][m2];
                    Hi[l][m2] = s;
                }
            for (int l = 0; l < PI; l++)
                for (int m2 = 0; m2 < PI; m2++) Ho[((size_t)j * PI + l) * PI + m2] = Hi[l][m2];
        }
    }
}

__global__ void fill_kernel(float *__restrict__ p, size_t n, float v)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static inline size_t rec_bytes(const gt_atoms *A, const gt_geom *G)
{
    return align_up((size_t)A->n * G->nA * sizeof(PairRec));
}
static inline size_t mom_bytes(const gt_atoms *A, const gt_geom *G)
{
    return align_up((size_t)A->n * G->nA * MOM_STRIDE * sizeof(float));
}

int launch_setup(const gt_atoms *A, const gt_geom *G, PairRec *recs, hipStream_t st)
{
    const long total = (long)A->n * G->nA;
    if (total == 0) return 0;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    if (G->dim == 3)
        launch_timed(GT_K_SETUP, setup_kernel<3>, dim3(blocks), dim3(256), 0, st, *A, *G, recs);
    else
        launch_timed(GT_K_SETUP, setup_kernel<2>, dim3(blocks), dim3(256), 0, st, *A, *G, recs);
    GT_HIP(hipGetLastError());
    return 0;
}

int launch_fill(float *p, size_t n, float v, hipStream_t st)
{
    if (n == 0) return 0;
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 65535);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, st, p, n, v);
    GT_HIP(hipGetLastError());
    return 0;
}

// Forward projection from precomputed records (used by the fitters too).
__global__ void negate_kernel(float *__restrict__ R, const float *__restrict__ sub, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        R[i] = 0.0f - sub[i];
}

// Forward projection from precomputed records (used by the fitters too);
// with sub != NULL the result is the residual projection - sub.
int launch_forward(const gt_atoms *A, const gt_geom *G, const PairRec *recs, float *R, hipStream_t st,
                   const float *sub)
{
    const size_t nout = (size_t)G->nA * G->n0 * G->n1;
    if (A->n == 0) {
        if (!sub) return launch_fill(R, nout, 0.0f, st);
        if (nout == 0) return 0;
        hipLaunchKernelGGL(negate_kernel, dim3((unsigned)std::min<size_t>((nout + 255) / 256, 65535)), dim3(256), 0,
                           st, R, sub, nout);
        GT_HIP(hipGetLastError());
        return 0;
    }
    if (G->dim == 3) {
        const int e0 = ceil_log2((G->n0 + TILE_ROWS - 1) / TILE_ROWS), e1 = ceil_log2((G->n1 + TILE_COLS - 1) / TILE_COLS);
        const unsigned long long blocks = (unsigned long long)G->nA << (e0 + e1);
        GT_CHECK_ARG(blocks < (1ull << 31), "forward: too many detector tiles (%d x %d, %d angles)", G->n0, G->n1,
                     G->nA);
        launch_timed(GT_K_FORWARD, forward3_kernel, dim3((unsigned)blocks), dim3(F3_WAVES * 64), 0, st, recs, A->n,
                     G->nA, G->n0, G->n1, G->p0, G->p1, R, sub, e0, e1, G->sorted);
    } else {
        const int tiles = (G->n0 + 255) / 256;
        launch_timed(GT_K_FORWARD, forward2_kernel, dim3(tiles, G->nA), dim3(256), 0, st, recs, A->n, G->n0, G->p0, R,
                     sub);
    }
    GT_HIP(hipGetLastError());
    return 0;
}

template <int D, int MODE>
static void launch_deriv_kernels(const gt_atoms *A, const gt_geom *G, const PairRec *recs, float *mom,
                                 const double *sumC2, const float *C, int want_hess, double *f, double *g, double *H,
                                 hipStream_t st)
{
    constexpr int K = D - 1;
    const long pairs = (long)A->n * G->nA;
    launch_timed(GT_K_MOMENTS, moments_kernel<K, MODE>, dim3((unsigned)((pairs + MOM_WAVES - 1) / MOM_WAVES)),
                 dim3(64 * MOM_WAVES), 0, st, recs,
                 A->n, G->nA, G->n0, G->n1, G->p0, G->p1, C, mom);
    launch_timed(GT_K_FINISH, finish_kernel<D, MODE>, dim3(A->n), dim3(64), 0, st, *A, *G, recs, mom, sumC2, want_hess,
                 f, g, H);
}

int launch_derivs(const gt_atoms *A, const gt_geom *G, const PairRec *recs, float *mom, double *sumC2,
                  const float *C, int mode, int want_hess, double *f, double *g, double *H, hipStream_t st)
{
    if (A->n == 0) return 0;
    GT_CHECK_ARG((long)A->n * G->nA / 4 < (1L << 31), "too many (atom, angle) pairs for one launch");
    if (mode == MODE_L2 && f) {
        hipLaunchKernelGGL(sumsq_kernel, dim3(G->nA), dim3(256), 0, st, C, (long)G->n0 * G->n1, sumC2);
    }
    const int D = G->dim;
#define GT_DISPATCH(DD, MM)                                                                             \
    if (D == DD && mode == MM)                                                                          \
        launch_deriv_kernels<DD, MM>(A, G, recs, mom, sumC2, C, want_hess, (mode == MODE_RESID) ? nullptr : f, g, \
                                     H, st);
    GT_DISPATCH(3, MODE_LINEAR)
    GT_DISPATCH(3, MODE_L2)
    GT_DISPATCH(3, MODE_RESID)
    GT_DISPATCH(2, MODE_LINEAR)
    GT_DISPATCH(2, MODE_L2)
    GT_DISPATCH(2, MODE_RESID)
#undef GT_DISPATCH
    GT_HIP(hipGetLastError());
    return 0;
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_forward_workspace(const gt_atoms *atoms, const gt_geom *geom)
{
    if (!atoms || !geom) return 0;
    return rec_bytes(atoms, geom);
}

extern "C" int gt_forward(const gt_atoms *atoms, const gt_geom *geom, float *R, void *ws, size_t ws_bytes,
                          void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_geom(geom, atoms->dim))) return rc;
    GT_CHECK_ARG(R != nullptr, "gt_forward: R is NULL");
    const size_t need = gt_forward_workspace(atoms, geom);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_forward: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "gt_forward: workspace must be 64-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    PairRec *recs = (PairRec *)ws;
    if ((rc = launch_setup(atoms, geom, recs, st))) return rc;
    return launch_forward(atoms, geom, recs, R, st, nullptr);
}

extern "C" size_t gt_derivs_workspace(const gt_atoms *atoms, const gt_geom *geom, int mode)
{
    (void)mode;
    if (!atoms || !geom) return 0;
    return rec_bytes(atoms, geom) + mom_bytes(atoms, geom) + align_up((size_t)geom->nA * sizeof(double));
}

extern "C" int gt_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *C, int mode, int order, double *f,
                         double *g, double *H, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_geom(geom, atoms->dim))) return rc;
    GT_CHECK_ARG(mode == GT_MODE_LINEAR || mode == GT_MODE_L2 || mode == GT_MODE_RESID,
                 "gt_derivs: mode must be 0 (linear), 1 (L2) or 2 (residual), got %d", mode);
    GT_CHECK_ARG(order == 1 || order == 2, "gt_derivs: order must be 1 or 2, got %d", order);
    GT_CHECK_ARG(C != nullptr || atoms->n == 0, "gt_derivs: C is NULL");
    GT_CHECK_ARG(g != nullptr || atoms->n == 0, "gt_derivs: g is NULL");
    const size_t need = gt_derivs_workspace(atoms, geom, mode);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_derivs: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "gt_derivs: workspace must be 64-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    char *w = (char *)ws;
    PairRec *recs = (PairRec *)w;
    float *mom = (float *)(w + rec_bytes(atoms, geom));
    double *sumC2 = (double *)(w + rec_bytes(atoms, geom) + mom_bytes(atoms, geom));
    if ((rc = launch_setup(atoms, geom, recs, st))) return rc;
    return launch_derivs(atoms, geom, recs, mom, sumC2, C, mode, order > 1, f, g, H, st);
}

extern "C" int gt_forward_derivs(const gt_atoms *atoms, const gt_geom *geom, const float *data, int order, float *R,
                                 double *f, double *g, double *H, void *ws, size_t ws_bytes, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_geom(geom, atoms->dim))) return rc;
    GT_CHECK_ARG(order == 1 || order == 2, "gt_forward_derivs: order must be 1 or 2, got %d", order);
    GT_CHECK_ARG(data != nullptr, "gt_forward_derivs: data is NULL");
    GT_CHECK_ARG(R != nullptr, "gt_forward_derivs: R is NULL");
    GT_CHECK_ARG(g != nullptr || atoms->n == 0, "gt_forward_derivs: g is NULL");
    GT_CHECK_ARG(R != data, "gt_forward_derivs: R must not alias data");
    const size_t need = gt_derivs_workspace(atoms, geom, GT_MODE_RESID);
    if (need > ws_bytes || (need > 0 && ws == nullptr)) {
        set_error("gt_forward_derivs: workspace of %zu bytes given, %zu needed", ws_bytes, need);
        return GT_EWORKSPACE;
    }
    GT_CHECK_ARG(((uintptr_t)ws & 63) == 0, "gt_forward_derivs: workspace must be 64-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    char *w = (char *)ws;
    PairRec *recs = (PairRec *)w;
    float *mom = (float *)(w + rec_bytes(atoms, geom));
    double *sumC2 = (double *)(w + rec_bytes(atoms, geom) + mom_bytes(atoms, geom));
    if ((rc = launch_setup(atoms, geom, recs, st))) return rc;
    if ((rc = launch_forward(atoms, geom, recs, R, st, data))) return rc;
    if (f) {
        hipLaunchKernelGGL(sumsq_kernel, dim3(geom->nA), dim3(256), 0, st, R, (long)geom->n0 * geom->n1, sumC2);
        hipLaunchKernelGGL(half_total_kernel, dim3(1), dim3(1), 0, st, sumC2, geom->nA, f);
        GT_HIP(hipGetLastError());
    }
    return launch_derivs(atoms, geom, recs, mom, sumC2, R, GT_MODE_RESID, order > 1, nullptr, g, H, st);
}
