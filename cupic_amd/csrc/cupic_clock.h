// cupic_clock.h -- the injection clock's FP64 recurrence in closed form.
//
// One step inserts n particles at times tin_0, ..., tin_{n-1} and leaves the
// clock at tin_n, where tin_{k+1} = fl(tin_k + d) (src/cc.cu:179, 247: one
// FP64 add per inserted particle, round to nearest even).  Evaluated one add
// after another that is a chain of n dependent adds (about 8,000 per step
// at 1024x512 cells, 200 ppc).  It need not be:
//
// Inside one binade [2^e, 2^(e+1)) the doubles are the multiples of
// u = 2^(e-52).  With tin_k a multiple of u, the exact sum tin_k + d rounds
// to tin_k + delta, delta = d rounded to the nearest multiple of u, as long
// as that sum stays at or below 2^(e+1) -- except when d lies exactly half
// way between two multiples of u, where round-to-even makes the result
// depend on tin_k.  So the recurrence is an arithmetic progression
// tin_k = t0 + j delta (j = 0, 1, ...), exact in FP64, until it leaves the
// binade; the add that leaves it is done as an ordinary FP64 add, and a new
// progression starts in the next binade.  A tie, a zero delta, a d of a
// binade's size or more, or a subnormal tin make the segment a single
// ordinary add.
//
// clock_segments cuts the n values into such segments; every value of a
// segment is t0 + j delta with j < len, and the clock after the last is
// returned.  The result equals the chain of adds bit for bit.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cupic {

struct ClockSeg {
    int k0;        // index of the segment's first value
    int len;       // values in the segment
    double t0;     // first value
    double delta;  // step inside the segment (0 for a single value)
};

__host__ __device__ inline double clock_pow2(int e)  // 2^e for a normal exponent
{
    return __builtin_bit_cast(double, (long long)(e + 1023) << 52);
}

// value j of a segment
__host__ __device__ inline double clock_value(const ClockSeg &s, int j)
{
    return s.t0 + (double)j * s.delta;
}

// Cuts tin_0 .. tin_{n-1} into at most maxseg segments and stores the clock
// tin_n in *tin_end.  Returns the number of segments, or -1 when maxseg is
// too small (the caller then continues with plain adds from segment maxseg's
// start: seg[maxseg - 1] is complete and *tin_end holds the first value not
// covered, at index seg[maxseg-1].k0 + seg[maxseg-1].len).
__host__ __device__ inline int clock_segments(double tin, double d, int n, ClockSeg *seg,
                                              int maxseg, double *tin_end)
{
    int ns = 0, k = 0;
    while (k < n) {
        if (ns == maxseg) {
            *tin_end = tin;
            return -1;
        }
        const long long bits = __builtin_bit_cast(long long, tin);
        const int ef = (int)((bits >> 52) & 0x7FF);
        // tin in [2^e, 2^(e+1)), e = ef - 1023; grid spacing u = 2^(e-52)
        bool plain = (bits <= 0) || ef == 0 || ef >= 0x7FE || !(d > 0.0);
        double u = 0.0, hi = 0.0, delta = 0.0;
        long long room = 0, step = 0;
        if (!plain) {
            const int e = ef - 1023;
            u = clock_pow2(e - 52);
            hi = clock_pow2(e + 1);
            // d in units of u: exact (a power-of-two scale), integer part and
            // remainder exact while d / u < 2^52
            const double q = d / u;
            if (!(q < 4503599627370496.0)) {
                plain = true;  // d of a binade's size or more: no progression
            } else {
                const double qf = floor(q);
                const double r = q - qf;
                if (r == 0.5) {
                    plain = true;  // a tie: round-to-even depends on tin_k
                } else {
                    step = (long long)qf + (r > 0.5 ? 1 : 0);
                    delta = (double)step * u;
                    room = (long long)((hi - tin) / u);  // exact: a multiple of u below 2^53 u
                    if (step == 0) plain = true;
                }
            }
        }
        if (plain) {
            seg[ns].k0 = k;
            seg[ns].len = 1;
            seg[ns].t0 = tin;
            seg[ns].delta = 0.0;
            ns++;
            tin = tin + d;
            k++;
            continue;
        }
        // values tin + j delta, j = 0..m, stay in the binade (the last may
        // equal 2^(e+1), which the add from tin + (m-1) delta also reaches)
        const long long m = room / step;
        const long long left = (long long)(n - k);
        const long long len = m + 1 < left ? m + 1 : left;
        seg[ns].k0 = k;
        seg[ns].len = (int)len;
        seg[ns].t0 = tin;
        seg[ns].delta = delta;
        ns++;
        const double last = tin + (double)(len - 1) * delta;
        // the next value: inside the progression while len <= m, else the
        // add that leaves the binade
        tin = len <= m ? tin + (double)len * delta : last + d;
        k += (int)len;
    }
    *tin_end = tin;
    return ns;
}

}  // namespace cupic
