// Host check of the culling predicates of the pixel kernels (gt_pixel.h):
// for random anisotropic 3D atoms on tilt-series angles,
//  * every detector pixel of a footprint box whose FP32 exponent passes the
//    cutoff (q >= kappa, evaluated with the kernels' expressions) lies inside
//    its column's row span (column_span, the moments walk);
//  * every 8 x 16 pixel rectangle holding such a pixel is kept by
//    support_meets_rect (the forward kernel's per-wave cull).
// Usage: span_check <atoms> <detector side> <window rows> <nonuniform 0|1>;
// prints "pairs in_support walked box rect_hits rect_kept" and exits 1 on the
// first violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "gt_pixel.h"

using namespace gt;

int main(int argc, char **argv)
{
    const int natoms = argc > 1 ? atoi(argv[1]) : 300;
    const int S = argc > 2 ? atoi(argv[2]) : 128;
    const int TROWS = argc > 3 ? atoi(argv[3]) : 512;
    const int nonuni = argc > 4 ? atoi(argv[4]) : 0;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    std::normal_distribution<double> N01(0.0, 1.0);
    std::vector<float> p(S);
    for (int i = 0; i < S; i++) {
        double v = -1.0 + (i + 0.5) * 2.0 / S;
        if (nonuni) v += 0.3 * (2.0 / S) * std::sin(7.0 * i);  // ascending, not a uniform grid
        p[i] = (float)v;
    }
    const float *det[2] = {p.data(), p.data()};
    const int nd[2] = {S, S};
    long pairs = 0, in_support = 0, walked = 0, box = 0, rect_hits = 0, rect_kept = 0;
    for (int j = 0; j < natoms; j++) {
        PairGeom<3> g;
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++) g.U[a][b] = 0.0;
        double rho[3], L[3][3] = {{0}};
        for (int a = 0; a < 3; a++) rho[a] = 0.02 + 0.08 * U01(rng);
        for (int a = 0; a < 3; a++) L[a][a] = 1.0 / rho[a];
        L[1][0] = 5.0 * N01(rng);
        L[2][0] = 5.0 * N01(rng);
        L[2][1] = 5.0 * N01(rng);
        for (int a = 0; a < 3; a++)
            for (int b = a; b < 3; b++) g.U[a][b] = L[b][a] / std::sqrt(60.0);  // U = L^T / sqrt 60
        for (int a = 0; a < 3; a++) g.x[a] = -0.7 + 1.4 * U01(rng);
        g.I = 1.0;
        for (int t = 0; t < 61; t++) {
            const double th = (-60.0 + 2.0 * t) * M_PI / 180.0;
            // tilt about y: ray direction, detector axes
            g.T[0] = -std::sin(th); g.T[1] = 0.0; g.T[2] = std::cos(th);
            g.G[0][0] = std::cos(th); g.G[0][1] = 0.0; g.G[0][2] = std::sin(th);
            g.G[1][0] = 0.0; g.G[1][1] = 1.0; g.G[1][2] = 0.0;
            PairRec rec;
            pair_setup<3>(g, det, nd, 1, rec);
            const int Hbox = rec.hi[0] - rec.lo[0] + 1, W = rec.hi[1] - rec.lo[1] + 1;
            if (Hbox <= 0 || W <= 0) continue;
            pairs++;
            box += (long)Hbox * W;
            // forward cull: rectangles of 8 rows x 16 columns on the detector grid
            const float4 f1 = make_float4(rec.c_lo[1], rec.q[0], rec.q[1], rec.q[2]);
            const float4 f2 = make_float4(rec.q[3], rec.q[4], rec.q[5], 0.0f);
            for (int r0 = (rec.lo[0] / 8) * 8; r0 <= rec.hi[0]; r0 += 8)
                for (int c0 = (rec.lo[1] / 16) * 16; c0 <= rec.hi[1]; c0 += 16) {
                    const int r1 = r0 + 7 < S - 1 ? r0 + 7 : S - 1, c1 = c0 + 15 < S - 1 ? c0 + 15 : S - 1;
                    bool any = false;
                    for (int r = r0; r <= r1 && !any; r++)
                        for (int c = c0; c <= c1 && !any; c++) {
                            const float dp0 = (p[r] - rec.c_hi[0]) - rec.c_lo[0];
                            const float dp1 = (p[c] - rec.c_hi[1]) - rec.c_lo[1];
                            any = pair_q(rec.q, dp0, dp1) >= (float)KAPPA;
                        }
                    const bool keep = support_meets_rect(
                        f1, f2, (p[r0] - rec.c_hi[0]) - rec.c_lo[0], (p[r1] - rec.c_hi[0]) - rec.c_lo[0],
                        (p[c0] - rec.c_hi[1]) - rec.c_lo[1], (p[c1] - rec.c_hi[1]) - rec.c_lo[1]);
                    rect_hits++;
                    rect_kept += keep;
                    if (any && !keep) {
                        printf("FAIL rectangle atom %d angle %d rows %d.. cols %d..\n", j, t, r0, c0);
                        return 1;
                    }
                }
            for (int w0 = 0; w0 < Hbox; w0 += TROWS) {
                const int H = Hbox - w0 < TROWS ? Hbox - w0 : TROWS;
                const int row0 = rec.lo[0] + w0;
                std::vector<float> x(H);
                for (int r = 0; r < H; r++) x[r] = (p[row0 + r] - rec.c_hi[0]) - rec.c_lo[0];
                // row_table's grid test
                float dx = 0.0f;
                if (H >= 2) {
                    dx = (x[H - 1] - x[0]) / (float)(H - 1);
                    bool ok = dx > 0.0f;
                    for (int r = 0; r < H; r++) ok = ok && std::fabs(x[r] - std::fmaf((float)r, dx, x[0])) <= 0.25f * dx;
                    if (!ok) dx = 0.0f;
                }
                for (int c = 0; c < W; c++) {
                    const int i1 = rec.lo[1] + c;
                    const float dp1 = (p[i1] - rec.c_hi[1]) - rec.c_lo[1];
                    const float rowlin = std::fmaf(rec.q[1], dp1, rec.q[3]);
                    const float colt = std::fmaf(std::fmaf(rec.q[2], dp1, rec.q[4]), dp1, rec.q[5]);
                    int rs, re;
                    column_span(rec, rowlin, colt, x[0], dx, span_inv_dx(dx), span_inv2a(rec), H, rs, re);
                    if (re >= rs) walked += re - rs + 1;
                    for (int r = 0; r < H; r++) {
                        const float q = std::fmaf(std::fmaf(rec.q[0], x[r], rowlin), x[r], colt);
                        if (q >= (float)KAPPA) {
                            in_support++;
                            if (r < rs || r > re) {
                                printf("FAIL atom %d angle %d row %d col %d span [%d, %d] q %.9g\n", j, t, w0 + r, c,
                                       rs, re, q);
                                return 1;
                            }
                        }
                    }
                }
            }
        }
    }
    printf("%ld %ld %ld %ld %ld %ld\n", pairs, in_support, walked, box, rect_hits, rect_kept);
    return 0;
}
