// Host check of the detector index searches of gt_pair.h (lower_idx,
// upper_idx) against std::lower_bound / std::upper_bound on ascending float
// and double vectors: uniform grids (cell centres and end-point grids, as the
// projection spaces build them), non-uniform and clustered vectors, repeated
// values, one- and two-entry vectors, and query values on the grid points,
// between them, beyond both ends, +-inf and NaN.  Prints the number of
// queries checked and exits non-zero at the first mismatch.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "gt_pair.h"

template <typename T> static int ref_lower(const std::vector<T> &p, double v)
{
    if (std::isnan(v)) return 0;
    return (int)(std::lower_bound(p.begin(), p.end(), v, [](T a, double x) { return (double)a < x; }) - p.begin());
}
template <typename T> static int ref_upper(const std::vector<T> &p, double v)
{
    if (std::isnan(v)) return -1;
    return (int)(std::upper_bound(p.begin(), p.end(), v, [](double x, T a) { return x < (double)a; }) - p.begin()) -
           1;
}

static long checked = 0;

template <typename T> static bool check(const std::vector<T> &p, std::mt19937_64 &rng)
{
    const int n = (int)p.size();
    std::vector<double> qs;
    const double lo = p.front(), hi = p.back(), span = hi - lo > 0 ? hi - lo : 1.0;
    for (int i = 0; i < n; i++) {
        const double x = p[i];
        qs.push_back(x);
        qs.push_back(std::nextafter(x, -INFINITY));
        qs.push_back(std::nextafter(x, INFINITY));
        qs.push_back((double)std::nextafter((T)x, (T)-INFINITY));
        qs.push_back((double)std::nextafter((T)x, (T)INFINITY));
        if (i + 1 < n) qs.push_back(0.5 * (x + (double)p[i + 1]));
    }
    std::uniform_real_distribution<double> u(lo - 0.3 * span, hi + 0.3 * span);
    for (int k = 0; k < 4 * n + 20; k++) qs.push_back(u(rng));
    qs.push_back(INFINITY);
    qs.push_back(-INFINITY);
    qs.push_back(std::numeric_limits<double>::quiet_NaN());
    qs.push_back(lo - 1e300);
    qs.push_back(hi + 1e300);
    for (double v : qs) {
        const int gl = gt::lower_idx(p.data(), n, v), rl = ref_lower(p, v);
        const int gu = gt::upper_idx(p.data(), n, v), ru = ref_upper(p, v);
        checked++;
        if (gl != rl || gu != ru) {
            std::printf("MISMATCH n=%d v=%.17g lower %d (want %d) upper %d (want %d)\n", n, v, gl, rl, gu, ru);
            return false;
        }
    }
    return true;
}

template <typename T> static bool run(std::mt19937_64 &rng)
{
    const int sizes[] = {1, 2, 3, 5, 64, 96, 127, 128, 192, 256, 1000};
    for (int n : sizes) {
        std::vector<T> p(n);
        // cell centres of [-1, 1] (the detector partitions) and end-point grids
        for (int i = 0; i < n; i++) p[i] = (T)(-1.0 + 2.0 / n * (i + 0.5));
        if (!check(p, rng)) return false;
        for (int i = 0; i < n; i++) p[i] = (T)(n > 1 ? -1.3 + 2.6 * i / (n - 1) : 0.25);
        if (!check(p, rng)) return false;
        // a shifted, scaled uniform grid
        for (int i = 0; i < n; i++) p[i] = (T)(3.7 + 0.013 * i);
        if (!check(p, rng)) return false;
        // non-uniform: sorted random, clustered (squares), with repeats
        std::uniform_real_distribution<double> u(-2.0, 2.0);
        for (int i = 0; i < n; i++) p[i] = (T)u(rng);
        std::sort(p.begin(), p.end());
        if (!check(p, rng)) return false;
        for (int i = 0; i < n; i++) p[i] = (T)((double)i * i / ((double)n * n));
        if (!check(p, rng)) return false;
        for (int i = 0; i < n; i++) p[i] = (T)(i / 3);
        if (!check(p, rng)) return false;
        for (int i = 0; i < n; i++) p[i] = (T)0.5;
        if (!check(p, rng)) return false;
    }
    return true;
}

int main()
{
    std::mt19937_64 rng(12345);
    if (!run<float>(rng) || !run<double>(rng)) return 1;
    std::printf("checked %ld queries\n", checked);
    return 0;
}
