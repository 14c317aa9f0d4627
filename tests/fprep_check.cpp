// Host check of vs_fprep.hpp: repeat_add(a, d, k) against k FP32 additions
// done one by one, bit for bit, on seeded random runs that include exact
// halfway steps, runs that cross binade edges and zero, fixed points and the
// ranges the raycast march uses. Prints the number of runs and mismatches.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>

#include "vs_fprep.hpp"

static float serial(float a, float d, int k) {
    for (int i = 0; i < k; ++i) a = a + d;
    return a;
}

int main(int argc, char **argv) {
    const long n = argc > 1 ? atol(argv[1]) : 200000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto logu = [&](double lo2, double hi2) { return std::exp2(lo2 + (hi2 - lo2) * U(rng)); };
    auto sgn = [&]() { return U(rng) < 0.5 ? -1.0 : 1.0; };
    long bad = 0, runs = 0, jumped = 0;
    for (long i = 0; i < n; ++i) {
        float a, d;
        int k;
        const int kind = (int)(i % 6);
        if (kind == 0) {  // march positions and steps
            a = (float)(-700.0 + 2900.0 * U(rng));
            d = (float)(sgn() * 2.7 * U(rng));
            k = (int)(U(rng) * 600);
        } else if (kind == 1) {  // march totals
            a = (float)(900.0 * U(rng) * U(rng));
            d = (float)(0.2 + 3.0 * U(rng));
            k = (int)(U(rng) * 600);
        } else if (kind == 2) {  // exact halfway steps: d = (m + 1/2) u of a's binade
            const int e = (int)(U(rng) * 12) - 2;
            a = (float)(sgn() * std::ldexp(1.0 + U(rng), e));
            const double u = std::ldexp(1.0, e - 23);
            d = (float)(sgn() * (std::floor(U(rng) * 5000.0) + 0.5) * u);
            k = (int)(U(rng) * 4000);
        } else if (kind == 3) {  // next to a binade edge, both directions
            const int e = (int)(U(rng) * 14) - 3;
            const double edge = std::ldexp(1.0, e);
            const double off = std::ldexp(1.0, e - 23) * std::floor(U(rng) * 64.0);
            const bool up = U(rng) < 0.5;
            a = (float)(sgn() * (up ? edge - off : edge + off));
            d = (float)((a > 0) == up ? 1.0 : -1.0) * (float)logu(-30, 2);
            k = (int)(U(rng) * 3000);
        } else if (kind == 4) {  // crossing zero, tiny and subnormal values
            a = (float)(sgn() * logu(-140, 4));
            d = (float)(-(a > 0 ? 1.0 : -1.0) * logu(-40, 3));
            k = (int)(U(rng) * 3000);
        } else {  // fixed points and steps far below the ulp
            a = (float)(sgn() * logu(-10, 14));
            d = (float)(sgn() * std::fabs(a) * logu(-40, -20));
            if (U(rng) < 0.1) d = 0.0f;
            k = (int)(U(rng) * 3000);
        }
        const float want = serial(a, d, k), got = vs::repeat_add(a, d, k);
        ++runs;
        if (vs::f32_bits(want) != vs::f32_bits(got)) {
            if (bad < 10) printf("mismatch a=%a d=%a k=%d want=%a got=%a\n", a, d, k, want, got);
            ++bad;
        }
        if (k > 3 && vs::binade_room(a + d, a + d + d, a + d + d + d) > 0) ++jumped;
    }
    printf("runs %ld jumped %ld mismatches %ld\n", runs, jumped, bad);
    return bad == 0 ? 0 : 1;
}
