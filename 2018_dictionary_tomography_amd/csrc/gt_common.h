// gt_common.h -- shared host/device helpers of the HIP library.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gausstomo.h"
#include "gt_pair.h"

namespace gt {

// Thread-local last error message (gt_last_error).
void set_error(const char *fmt, ...);

// Optional per-launch event timing (gt_profile_enable / gt_profile_read).
void prof_begin(hipStream_t st, int id);
void prof_end(hipStream_t st, int id);
// The same for a launch timed by its own dispatch: events a, b for
// hipExtLaunchKernelGGL, which records the kernel's start and end in the
// kernel's packet (no packets between launches); false, and both null, when
// the kernel id is not being timed.
bool prof_ext_begin(int id, hipEvent_t *a, hipEvent_t *b);
void prof_ext_end(int id, hipEvent_t a, hipEvent_t b);

template <typename F, typename... Args>
inline void launch_timed(int id, F kernel, dim3 grid, dim3 block, unsigned shmem, hipStream_t st, Args... args)
{
    hipEvent_t a = nullptr, b = nullptr;
    const bool timed = prof_ext_begin(id, &a, &b);
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, st, a, b, 0u, args...);
    if (timed) prof_ext_end(id, a, b);
}

#define GT_CHECK_ARG(cond, ...)          \
    do {                                 \
        if (!(cond)) {                   \
            ::gt::set_error(__VA_ARGS__); \
            return GT_EINVAL;            \
        }                                \
    } while (0)

#define GT_HIP(call)                                                          \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            ::gt::set_error("%s failed: %s", #call, hipGetErrorString(e_));  \
            return GT_EHIP;                                                   \
        }                                                                     \
    } while (0)

constexpr size_t ALIGN = 256;
__host__ __device__ inline size_t align_up(size_t v) { return (v + ALIGN - 1) / ALIGN * ALIGN; }

// Centre-first numbering of tile grids.  Tiles at the centre of a detector,
// volume or k grid meet the most atoms; kernels whose grids hold more blocks
// than fit at once number their blocks so that block b takes the tile whose
// centre-out ranks (t0, t1, t2) have Morton index b (bits dealt to the axes in
// turn while each has bits left; grid of 2^e0 x 2^e1 x 2^e2 ranks).  Ranks
// beyond an axis' tile count are empty blocks.

// t-th index of 0..R-1 from the centre outwards: c, c+1, c-1, c+2, ... with c = (R-1)/2
__device__ inline int centre_out(int t, int R)
{
    const int c = (R - 1) >> 1;
    return (t & 1) ? c + ((t + 1) >> 1) : c - (t >> 1);
}

// Morton index m -> ranks t[0..2] of a 2^e[0] x 2^e[1] x 2^e[2] grid
__device__ inline void morton_ranks(unsigned m, const int *e, int *t)
{
    t[0] = t[1] = t[2] = 0;
    for (int bit = 0; m != 0u && bit < 31; bit++)
#pragma unroll
        for (int d = 0; d < 3; d++)
            if (bit < e[d]) {
                t[d] |= (int)(m & 1u) << bit;
                m >>= 1;
            }
}

// smallest e with 2^e >= v (v >= 1)
inline int ceil_log2(long v)
{
    int e = 0;
    while ((1L << e) < v) e++;
    return e;
}

// T in global memory.  Loads and stores through a gmem<T> * are global_*
// instructions (scalar base plus 32-bit offset, counted by the vector-memory
// counter only).  Through a plain pointer the compiler cannot trace to a
// kernel argument (one read from a struct in memory) they would be flat
// ones: 64-bit address arithmetic per access, and counted by the LDS counter
// too, so that every wait for an LDS read also waits for them.
template <typename T> using gmem = __attribute__((address_space(1))) T;

// Raw buffer resource over a read-only array (gfx9 word 3; no bounds beyond
// 32-bit offsets): buffer loads add a per-lane offset, a wave-uniform scalar
// offset and the base in the load itself, so a group of loads whose
// addresses differ by uniform amounts needs one vector offset.
__device__ inline __amdgpu_buffer_rsrc_t buffer_rsrc(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, -1, 0x00020000);
}
__device__ inline float buffer_load_f32(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// LDS written by some lanes of a wave, then read by others: order the wave's
// LDS traffic (the lanes of a wave execute together).
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// sin(2 pi t) and cos(2 pi t) of a phase t in turns, |t| <= 256: the
// hardware sine and cosine, whose argument is in turns (quarter-rate, like
// exp2); measured maximum absolute error 1.25e-7 for |t| <= 1.5 on gfx950.
__device__ inline void ft_sincos_turns(float t, float &sn, float &cs)
{
    sn = __builtin_amdgcn_sinf(t);
    cs = __builtin_amdgcn_cosf(t);
}

int check_atoms(const gt_atoms *a);
int check_geom(const gt_geom *g, int dim);

// Atom j into a PairGeom (iso: U = rho * Id).
template <int D> __device__ __host__ inline void load_atom(const gt_atoms &A, int j, PairGeom<D> &g)
{
    constexpr int NR = Dims<D>::NR;
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) g.U[i][k] = 0.0;
    if (A.iso) {
        double rho = A.r[j];
        for (int i = 0; i < D; i++) g.U[i][i] = rho;
    } else {
        // unrolled: constant indices keep U in registers
#pragma unroll
        for (int k = 0; k < NR; k++) g.U[slot_row(D, k)][slot_col(D, k)] = A.r[(size_t)j * NR + k];
    }
    for (int i = 0; i < 3; i++) g.x[i] = (i < D) ? (double)A.x[(size_t)j * D + i] : 0.0;
    g.I = A.I[j];
}

// Orientation a into a PairGeom: T and the detector columns G.
template <int D> __device__ __host__ inline void load_angle(const gt_geom &G, int a, PairGeom<D> &g)
{
    for (int i = 0; i < 3; i++) {
        g.T[i] = 0.0;
        g.G[0][i] = 0.0;
        g.G[1][i] = 0.0;
    }
    const int L = G.lenT;
    // constant indices (a loop to the run-time L would index T and G[0]
    // dynamically and put them in scratch memory)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        if (i < L) {
            g.T[i] = G.T[(size_t)a * L + i];
            g.G[0][i] = G.w0[(size_t)a * L + i];
        }
    }
    if (D == 3) {
        if (L == 3) {
            for (int i = 0; i < 3; i++) g.G[1][i] = G.w1[(size_t)a * 3 + i];
        } else {
            g.G[1][2] = 1.0;  // 2.5D: y = (p0 w0[0], p0 w0[1], p1), NewtonBases.py:684-688
        }
    }
}

// Wave64 helpers.
__device__ inline int lane_id() { return threadIdx.x & 63; }

}  // namespace gt
