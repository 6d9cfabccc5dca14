// The value of a running f64 sum after n additions of one step, as a loop that adds the step n times
// forms it: s_0 = s, s_(k+1) = RN(s_k + t).  The reference walks its lens-path rows this way,
// local += world_to_local[:, 0] once per output pixel (remapU8RowsNoTranslationLens,
// src/remap.zig:404-437), so pixel x of a row sees x roundings.  A thread here needs s_x for its own
// x without the x - 1 values before it.
//
// Inside one binade [2^e, 2^(e+1)) every double is an integer multiple of u = 2^(e-52), so adding t
// to such a value rounds the exact sum to the nearest multiple of u: s_(k+1) = s_k + D u, where D is t/u
// rounded to an integer (ties to even: S_k + D even, with S_k = s_k / u).  As long as the results stay
// inside the binade the sum is therefore arithmetic, and s_(k+j) = (S_k + j D) u exactly.  The function
// takes one real addition, then the longest such run that stays clear of both ends of the binade,
// and repeats: a row crosses a few binades on each side of zero, so a handful of rounds reach any x.
// Host and device compile it alike (the CPU test checks it against the loop itself).
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define FG_HD __host__ __device__
#else
#define FG_HD
#endif

namespace fg {

FG_HD inline double seq_add(double s, double t, int n) {
    const double two52 = 4503599627370496.0, two53 = 9007199254740992.0;
    while (n > 0) {
        s = s + t;  // one addition, exactly as the loop makes it
        n--;
        if (n == 0 || t == 0.0) return s;
        if (s == 0.0) continue;
        // the run from s inside the binade of |s|, in units of its ulp
        const double sg = s < 0.0 ? -1.0 : 1.0;
        const double a = s * sg;
        const int e = ilogb(a);                // a in [2^e, 2^(e+1))
        const double S = ldexp(a, 52 - e);     // an integer in [2^52, 2^53)
        const double q = ldexp(t * sg, 52 - e);  // the step as it moves |s|, in ulps
        if (!(fabs(q) < two52)) continue;      // one step leaves the binade: single additions
        const double qf = floor(q), fr = q - qf;
        double D;
        if (fr == 0.5) {
            // a tie every step: the result is the even neighbour; from an odd S the first
            // addition (a real one, next round) makes it even, and an even D keeps it even
            if (fmod(S, 2.0) != 0.0) continue;
            D = fmod(qf, 2.0) == 0.0 ? qf : qf + 1.0;
        } else {
            D = fr < 0.5 ? qf : qf + 1.0;
        }
        if (D == 0.0) return s;  // |t| below half an ulp of |s|: every further addition returns s
        // steps j = 1..m with S + j D in [2^52 + 1, 2^53 - 2], whose exact sums S + j D + (q - D)
        // lie inside the binade, so each rounds to the grid of u
        const double A = D > 0.0 ? (two53 - 2.0) - S : S - (two52 + 1.0);
        const double ad = fabs(D);
        double m = A >= ad ? floor(A / ad) : 0.0;
        if (m * ad > A) m -= 1.0;  // the quotient rounded up to an integer
        if (m > (double)n) m = (double)n;
        if (m <= 0.0) continue;
        s = sg * ldexp(S + m * D, e - 52);
        n -= (int)m;
    }
    return s;
}

}  // namespace fg
