// gt_reg.hip -- batched per-atom regulariser terms (gt_reg_eval).
//
// The terms the fitter kernels add per atom update, evaluated for every atom
// of a set: what reg(atoms), reg.grad(atoms) and reg.hess(atom) of
// GaussDictCode/regularisation.py:22-442 return, one thread per atom.
#include "gt_common.h"
#include "gt_reg.h"

namespace gt {

struct RegParams {
    double p[GT_REG_NPARAM];
};

template <int D>
__global__ __launch_bounds__(256) void reg_kernel(gt_atoms A, int kind, RegParams prm, double *__restrict__ val,
                                                  double *__restrict__ grad, double *__restrict__ hess)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= A.n) return;
    constexpr int NR = D * (D + 1) / 2;
    const int nr = A.iso ? 1 : NR;
    const int PI = 1 + D + nr;
    double th[1 + D + NR];
    th[0] = A.I[j];
    for (int i = 0; i < D; i++) th[1 + i] = A.x[(size_t)j * D + i];
    for (int k = 0; k < nr; k++) th[1 + D + k] = A.r[(size_t)j * nr + k];
    double g[1 + D + NR], H[(1 + D + NR) * (1 + D + NR)];
    const double v = reg_eval<D>(kind, prm.p, A.iso, th, grad ? g : nullptr, H);
    val[j] = v;
    if (grad) {
        for (int i = 0; i < PI; i++) grad[(size_t)j * PI + i] = g[i];
        if (hess)
            for (int i = 0; i < PI * PI; i++) hess[(size_t)j * PI * PI + i] = H[i];
    }
}

}  // namespace gt

using namespace gt;

extern "C" int gt_reg_eval(const gt_atoms *atoms, int kind, const double *params, double *val, double *grad,
                           double *hess, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms))) return rc;
    GT_CHECK_ARG(kind >= GT_REG_NULL && kind <= GT_REG_ISO, "gt_reg_eval: unknown regulariser kind %d", kind);
    GT_CHECK_ARG(params != nullptr || kind == GT_REG_NULL, "gt_reg_eval: params is NULL");
    GT_CHECK_ARG(val != nullptr || atoms->n == 0, "gt_reg_eval: val is NULL");
    GT_CHECK_ARG(hess == nullptr || grad != nullptr, "gt_reg_eval: hess needs grad");
    if (atoms->n == 0) return 0;
    RegParams prm;
    for (int i = 0; i < GT_REG_NPARAM; i++) prm.p[i] = params ? params[i] : 0.0;
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid((atoms->n + 255) / 256), block(256);
    if (atoms->dim == 3)
        hipLaunchKernelGGL(reg_kernel<3>, grid, block, 0, st, *atoms, kind, prm, val, grad, hess);
    else
        hipLaunchKernelGGL(reg_kernel<2>, grid, block, 0, st, *atoms, kind, prm, val, grad, hess);
    GT_HIP(hipGetLastError());
    return 0;
}
