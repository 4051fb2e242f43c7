// gt_noalloc.hip -- the forward + derivative sweep with no allocation.
//
// gt_forward_derivs takes its workspace (per-(atom, angle) records, pixel
// moments, per-angle FP64 sums) from the caller.  Callers that cannot
// allocate device memory at all (code under HIP graph capture, serving loops,
// harnesses that forbid hipMalloc) use gt_forward_derivs_noalloc instead: the
// same kernels, with the workspace in `scratch`, a file-scope device array of
// GT_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds until the
// process ends.
//
// `scratch` names a host-side stand-in in host code; its device address is
// taken once per device with hipGetSymbolAddress (with that device current,
// on the device's first call) and kept per device: only that address is
// ever handed to kernels.  Nothing in the workspace has to start at zero
// (every kernel writes what it later reads), so no clearing is done.  Calls
// on one GPU share its array: calls that overlap on one GPU (different
// streams, or different host threads) race on it.
#include <mutex>

#include "gt_common.h"

namespace gt {

static __device__ __align__(256) unsigned char scratch[GT_NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;

// Device address of `scratch` on the current device, taken on its first use.
static int scratch_address(void **out)
{
    static std::mutex mu;
    static void *addr[NOALLOC_MAX_DEVICES];
    int dev = -1;
    GT_HIP(hipGetDevice(&dev));
    GT_CHECK_ARG(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "gt_forward_derivs_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (addr[dev] == nullptr) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr) {
            set_error("gt_forward_derivs_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return GT_EHIP;
        }
        addr[dev] = p;
    }
    *out = addr[dev];
    return 0;
}

}  // namespace gt

using namespace gt;

extern "C" size_t gt_noalloc_scratch_bytes(void) { return GT_NOALLOC_SCRATCH_BYTES; }

extern "C" int gt_forward_derivs_noalloc(const gt_atoms *atoms, const gt_geom *geom, const float *data, int order,
                                         float *R, double *f, double *g, double *H, void *stream)
{
    int rc;
    if ((rc = check_atoms(atoms)) || (rc = check_geom(geom, atoms->dim))) return rc;
    const size_t need = gt_derivs_workspace(atoms, geom, GT_MODE_RESID);
    if (need > GT_NOALLOC_SCRATCH_BYTES) {
        set_error("gt_forward_derivs_noalloc: %d atoms x %d angles need %zu bytes of scratch, above the %zu-byte "
                  "limit; call gt_forward_derivs with a workspace",
                  atoms->n, geom->nA, need, (size_t)GT_NOALLOC_SCRATCH_BYTES);
        return GT_EWORKSPACE;
    }
    void *ws = nullptr;
    if ((rc = scratch_address(&ws))) return rc;
    return gt_forward_derivs(atoms, geom, data, order, R, f, g, H, ws, GT_NOALLOC_SCRATCH_BYTES, stream);
}
