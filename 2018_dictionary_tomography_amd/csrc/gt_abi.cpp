// gt_abi.cpp -- error reporting and argument validation of the C ABI.
#include <stdarg.h>
#include <stdio.h>

#include "gt_common.h"

namespace gt {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_atoms(const gt_atoms *a)
{
    GT_CHECK_ARG(a != nullptr, "atoms is NULL");
    GT_CHECK_ARG(a->dim == 2 || a->dim == 3, "atoms.dim must be 2 or 3, got %d", a->dim);
    GT_CHECK_ARG(a->n >= 0, "atoms.n must be >= 0, got %d", a->n);
    GT_CHECK_ARG(a->iso == 0 || a->iso == 1, "atoms.iso must be 0 or 1, got %d", a->iso);
    if (a->n > 0) GT_CHECK_ARG(a->I && a->x && a->r, "atoms.I, atoms.x and atoms.r must be non-NULL");
    return 0;
}

int check_geom(const gt_geom *g, int dim)
{
    GT_CHECK_ARG(g != nullptr, "geometry is NULL");
    GT_CHECK_ARG(g->dim == dim, "geometry.dim (%d) does not match atoms.dim (%d)", g->dim, dim);
    if (dim == 2)
        GT_CHECK_ARG(g->lenT == 2, "2D geometry needs lenT == 2, got %d", g->lenT);
    else
        GT_CHECK_ARG(g->lenT == 2 || g->lenT == 3, "3D geometry needs lenT 2 or 3, got %d", g->lenT);
    GT_CHECK_ARG(g->nA >= 1 && g->nA <= 65535, "geometry.nA must be in [1, 65535], got %d", g->nA);
    GT_CHECK_ARG(g->n0 >= 1 && g->n0 < (1 << 24), "geometry.n0 out of range: %d", g->n0);
    if (dim == 2)
        GT_CHECK_ARG(g->n1 == 1, "2D geometry needs n1 == 1, got %d", g->n1);
    else
        GT_CHECK_ARG(g->n1 >= 1 && g->n1 < (1 << 24), "geometry.n1 out of range: %d", g->n1);
    GT_CHECK_ARG(g->T && g->w0 && g->p0, "geometry.T, w0 and p0 must be non-NULL");
    if (dim == 3) GT_CHECK_ARG(g->p1 != nullptr, "3D geometry needs p1");
    if (dim == 3 && g->lenT == 3) GT_CHECK_ARG(g->w1 != nullptr, "3D geometry with lenT 3 needs w1");
    GT_CHECK_ARG(g->sorted == 0 || g->sorted == 1, "geometry.sorted must be 0 or 1");
    return 0;
}

}  // namespace gt

extern "C" const char *gt_last_error(void) { return gt::g_err; }

extern "C" const char *gt_version(void) { return "gausstomo-mi355x 0.1 (gfx950)"; }
