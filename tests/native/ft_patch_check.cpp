// Host check of the Fourier grid kernel's patch test (gt_ftpatch.h): for
// random anisotropic and isotropic FT-space atoms (the 3D workload's draw
// mapped to FT parameters: U = L^T, scaled as FtRec stores it) on the integer
// grid [-64, 64)^3, every 8 x 8 (k1, k2) patch at every k0 that holds a
// sample whose FP32 exponent passes the cutoff (the kernel's expressions) must
// be kept by ft_patch_min < cut (1 + 1e-4).  Usage: ft_patch_check <atoms>
// <iso 0|1>; prints "patches kept with_support" and exits 1 on a violation.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "gt_ftpatch.h"

using namespace gt;

int main(int argc, char **argv)
{
    const int natoms = argc > 1 ? atoi(argv[1]) : 20, iso = argc > 2 ? atoi(argv[2]) : 0;
    const double s = 0.8493218002880191;  // sqrt(log2(e) / 2), as FtRec.u
    const float cut = (float)(30.0 * s * s), keep_cut = cut * (1.0f + 1e-4f);
    std::mt19937_64 rng(99);
    std::uniform_real_distribution<double> U01(0.0, 1.0);
    std::normal_distribution<double> N01(0.0, 1.0);
    long patches = 0, kept = 0, with_support = 0;
    for (int j = 0; j < natoms; j++) {
        FtL r = {};
        r.amp = 1.0f;
        if (iso) {
            r.iso = 1.0f;
            r.u[0] = (float)(s / (2.0 + 20.0 * U01(rng)));  // u = sqrt(log2e/2) / r, support radius sqrt(30) r
        } else {
            double L[3][3] = {{0}};
            for (int a = 0; a < 3; a++) L[a][a] = 1.0 / (0.02 + 0.08 * U01(rng));
            L[1][0] = 5.0 * N01(rng);
            L[2][0] = 5.0 * N01(rng);
            L[2][1] = 5.0 * N01(rng);
            // FT parameters: U~ = 2 pi U^{-1} with U = L^T (upper); slots [00, 11, 22, 01, 12, 02]
            double Uu[3][3] = {{0}}, Ui[3][3] = {{0}};
            for (int a = 0; a < 3; a++)
                for (int b = a; b < 3; b++) Uu[a][b] = L[b][a];
            for (int c = 2; c >= 0; c--)
                for (int col = 0; col < 3; col++) {
                    double v = (c == col) ? 1.0 : 0.0;
                    for (int k = c + 1; k < 3; k++) v -= Uu[c][k] * Ui[k][col];
                    Ui[c][col] = v / Uu[c][c];
                }
            const int R[6] = {0, 1, 2, 0, 1, 0}, C[6] = {0, 1, 2, 1, 2, 2};
            for (int k = 0; k < 6; k++) r.u[k] = (float)(s * 2.0 * M_PI * Ui[R[k]][C[k]]);
        }
        for (int k0 = -64; k0 < 64; k0++)
            for (int a1 = -64; a1 < 64; a1 += 8)
                for (int a2 = -64; a2 < 64; a2 += 8) {
                    bool any = false;
                    for (int d1 = 0; d1 < 8 && !any; d1++)
                        for (int d2 = 0; d2 < 8 && !any; d2++) {
                            const float q0 = (float)k0, q1 = (float)(a1 + d1), q2 = (float)(a2 + d2);
                            float nq;
                            if (iso) {
                                nq = std::fmaf(q1, q1, std::fmaf(q0, q0, q2 * q2)) * (r.u[0] * r.u[0]);
                            } else {
                                const float v0 = r.u[0] * q0;
                                const float v1 = std::fmaf(r.u[1], q1, r.u[3] * q0);
                                const float v2 = std::fmaf(r.u[4], q1, std::fmaf(r.u[5], q0, r.u[2] * q2));
                                nq = std::fmaf(v0, v0, std::fmaf(v1, v1, v2 * v2));
                            }
                            any = nq < cut;
                        }
                    const bool keep = ft_patch_min(r, (float)k0, (float)a1, (float)(a1 + 7), (float)a2,
                                                   (float)(a2 + 7)) < keep_cut;
                    patches++;
                    kept += keep;
                    with_support += any;
                    if (any && !keep) {
                        printf("FAIL atom %d k0 %d patch (%d, %d)\n", j, k0, a1, a2);
                        return 1;
                    }
                }
    }
    printf("%ld %ld %ld\n", patches, kept, with_support);
    return 0;
}
