// Host-side check that interp_series_short (the fused kernel's evaluation of
// short series) returns interp_series's value, bit for bit, on random series:
// sorted, with repeated times, unsorted, with NaN times and values, 0..6 points.
#include "af_device.hpp"
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <random>
#include <cmath>
int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<double> u(-10, 10);
    long bad = 0, n_cases = 0;
    for (int it = 0; it < 300000; it++) {
        int n = (int)(g() % 7);  // 0..6
        double t[8], v[8];
        int mode = (int)(g() % 4);
        for (int i = 0; i < n; i++) { t[i] = std::floor(u(g)); v[i] = u(g); }
        if (mode == 0) std::sort(t, t + n);
        if (mode == 1 && n > 1) { std::sort(t, t + n); t[g() % n] = t[0]; std::sort(t, t + n); }
        if (mode == 3 && n > 0 && g() % 8 == 0) t[g() % n] = NAN;
        if (n > 0 && g() % 16 == 0) v[g() % n] = NAN;
        double q = (g() % 3 == 0) ? t[n > 0 ? g() % n : 0] : std::floor(u(g) * 4) / 4;
        if (g() % 50 == 0) q = NAN;
        const double* pt = t; const double* pv = v;
        double a = af::interp_series(pt, pv, n, 1.25, q);
        double b = af::interp_series_short(pt, pv, n, 1.25, q);
        n_cases++;
        if (std::memcmp(&a, &b, 8) != 0 && !(std::isnan(a) && std::isnan(b))) {
            if (bad < 5) printf("mismatch n=%d q=%g a=%.17g b=%.17g\n", n, q, a, b);
            bad++;
        }
    }
    printf("cases %ld mismatches %ld\n", n_cases, bad);
    return bad != 0;
}
