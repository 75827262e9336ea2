// Host check of the closed-form injection clock (cupic_amd/csrc/cupic_clock.h)
// against the chain of FP64 adds it replaces, tin_{k+1} = tin_k + d, bit for
// bit: random clocks and intervals, binade crossings, ties (d half way
// between two multiples of the grid spacing), intervals of a binade's size
// or more, subnormal clocks, and segment lists too short for the values
// (the kernel's plain-add fallback).  Prints "cases C values V mismatches M".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "cupic_clock.h"

using cupic::ClockSeg;

// what k_clock computes: segments, then plain adds where the list ran out
static double closed_form(double tin, double d, int n, int maxseg, std::vector<double> &out)
{
    std::vector<ClockSeg> seg(maxseg);
    double end;
    int ns = cupic::clock_segments(tin, d, n, seg.data(), maxseg, &end);
    out.assign(n, 0.0);
    int covered = 0;
    for (int q = 0; q < (ns < 0 ? maxseg : ns); q++) {
        for (int j = 0; j < seg[q].len; j++) out[seg[q].k0 + j] = cupic::clock_value(seg[q], j);
        covered = seg[q].k0 + seg[q].len;
    }
    if (ns < 0) {
        for (int k = covered; k < n; k++) {
            out[k] = end;
            end += d;
        }
    }
    return end;
}

static double chain(double tin, double d, int n, std::vector<double> &out)
{
    out.assign(n, 0.0);
    for (int k = 0; k < n; k++) {
        out[k] = tin;
        tin += d;
    }
    return tin;
}

int main()
{
    std::mt19937_64 rng(20261203);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    long long cases = 0, values = 0, bad = 0;
    std::vector<double> a, b;
    auto check = [&](double tin, double d, int n, int maxseg) {
        double ea = closed_form(tin, d, n, maxseg, a);
        double eb = chain(tin, d, n, b);
        cases++;
        values += n;
        bool ok = ea == eb || (std::isnan(ea) && std::isnan(eb));
        for (int k = 0; k < n && ok; k++) ok = a[k] == b[k];
        if (!ok) {
            if (bad < 5)
                std::printf("mismatch tin %.17g d %.17g n %d maxseg %d\n", tin, d, n, maxseg);
            bad++;
        }
    };
    for (int it = 0; it < 20000; it++) {
        double tin = std::pow(10.0, -3.0 + 6.0 * U(rng));
        double d = std::pow(10.0, -9.0 + 8.5 * U(rng));
        int n = (int)(U(rng) * 3000.0);
        check(tin, d, n, it % 3 == 0 ? 2 : 64);
    }
    // the clock just below a power of two: one or more binade crossings
    for (int it = 0; it < 5000; it++) {
        int e = (int)(U(rng) * 20.0) - 10;
        double d = std::pow(10.0, -7.0 + 5.0 * U(rng)) * std::ldexp(1.0, e);
        double tin = std::ldexp(1.0, e) - d * (U(rng) * 50.0);
        check(tin, d, (int)(U(rng) * 2000.0) + 1, 64);
        check(tin, d, (int)(U(rng) * 2000.0) + 1, 3);
    }
    // ties: d an odd multiple of half the grid spacing of tin's binade
    for (int it = 0; it < 5000; it++) {
        int e = (int)(U(rng) * 16.0) - 8;
        double u = std::ldexp(1.0, e - 52);
        double tin = std::ldexp(1.0 + U(rng), e);
        tin = std::floor(tin / u) * u;
        double d = (std::floor(U(rng) * 1e9) + 0.5) * u;
        check(tin, d, (int)(U(rng) * 300.0) + 1, 64);
        check(tin, d, (int)(U(rng) * 300.0) + 1, 5);
    }
    // intervals of a binade's size or more, subnormal and zero clocks
    for (int it = 0; it < 2000; it++) {
        double tin = std::pow(10.0, -2.0 + 4.0 * U(rng));
        check(tin, tin * (0.5 + 3.0 * U(rng)), (int)(U(rng) * 100.0), 64);
    }
    check(1e-310, 3e-311, 500, 64);
    check(0.0, 1e-3, 500, 64);
    check(5e-324, 5e-324, 100, 64);
    check(1.0, 1e-3, 0, 64);
    // the production clock: t = 0.05 .. 250, dtin_e of 1024x512 cells, 200 ppc
    const double dtin_e = std::sqrt(2.0 * 3.1415926535897932) / (1600.0 * 512.0 * 0.5);
    for (int step = 1; step <= 5000; step += 7) {
        double t = 0.0;
        for (int s = 0; s < step; s++) t += 0.05;
        check(t + dtin_e * U(rng), dtin_e, 8200, 64);
    }
    std::printf("cases %lld values %lld mismatches %lld\n", cases, values, bad);
    return bad != 0;
}
