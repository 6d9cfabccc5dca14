// Checks fg::seq_add (focus-stacker_amd/csrc/fg_seqsum.h) against the loop it stands for: n plain f64
// additions of one step, the way the reference walks a lens-path row (src/remap.zig:404-437).  Test
// infrastructure only.  Cases: the rows of real lens poses (start of the row and step from
// world_to_local as the reference forms them, every column of rows up to 8000 wide), random starts
// and steps over many magnitudes and both signs, runs across zero, steps that tie at every addition
// (the ties-to-even rule), runs that reach either end of a binade, steps below half an ulp, and
// powers of two.  Prints the first mismatches and exits 1 if there are any.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <random>
#include <vector>

#include "../../focus-stacker_amd/csrc/fg_seqsum.h"

static long g_cases = 0, g_bad = 0;

static bool same(double a, double b) { return memcmp(&a, &b, sizeof a) == 0; }

// every n in [0, len] of one row against the running loop
static void check_row(double s0, double t, int len) {
    double s = s0;
    for (int n = 0; n <= len; n++) {
        const double r = fg::seq_add(s0, t, n);
        g_cases++;
        if (!same(r, s)) {
            if (g_bad < 12)
                printf("mismatch: s0=%.17g t=%.17g n=%d loop=%.17g seq_add=%.17g\n", s0, t, n, s, r);
            g_bad++;
        }
        s = s + t;
    }
}

// selected n only (long rows)
static void check_sparse(double s0, double t, int len, std::mt19937_64& rng) {
    std::vector<double> v(len + 1);
    v[0] = s0;
    for (int n = 1; n <= len; n++) v[n] = v[n - 1] + t;
    std::uniform_int_distribution<int> pick(0, len);
    for (int i = 0; i < 64; i++) {
        const int n = pick(rng);
        const double r = fg::seq_add(s0, t, n);
        g_cases++;
        if (!same(r, v[n])) {
            if (g_bad < 12)
                printf("mismatch: s0=%.17g t=%.17g n=%d loop=%.17g seq_add=%.17g\n", s0, t, n, v[n], r);
            g_bad++;
        }
    }
}

int main() {
    std::mt19937_64 rng(20261204);
    std::uniform_real_distribution<double> uni(0.0, 1.0);

    // lens-pose rows: local = w2l (x - cx, y - cy, -f), x stepping from the ROI's left edge
    for (int p = 0; p < 40; p++) {
        const double yaw = (uni(rng) - 0.5) * 0.02, pitch = (uni(rng) - 0.5) * 0.02, roll = (uni(rng) - 0.5) * 0.02;
        const double w = 800 + 7200 * uni(rng), h = 600 + 4800 * uni(rng);
        const double f = (w * 0.5) / tan(0.5 * (40.0 + 20.0 * uni(rng)) * M_PI / 180.0);
        const double cy = cos(yaw), sy = sin(yaw), cp = cos(pitch), sp = sin(pitch), cr = cos(roll), sr = sin(roll);
        const double m[3][3] = {{cr * cy + sr * sp * sy, sr * cp, -cr * sy + sr * sp * cy},
                                {-sr * cy + cr * sp * sy, cr * cp, sr * sy + cr * sp * cy},
                                {cp * sy, -sp, cp * cy}};
        const int left = (int)(20 * uni(rng));
        const double bx = (double)left - (w - 1.0) * 0.5;
        for (int r = 0; r < 6; r++) {
            const double py = (double)(int)(h * uni(rng)) - (h - 1.0) * 0.5;
            for (int i = 0; i < 3; i++) {
                const double s0 = m[i][0] * bx + m[i][1] * py + m[i][2] * -f;
                check_row(s0, m[i][0], (int)w);
            }
        }
    }
    // random starts and steps over many magnitudes, both signs, runs across zero
    for (int c = 0; c < 3000; c++) {
        const double s0 = (uni(rng) < 0.5 ? -1.0 : 1.0) * ldexp(uni(rng) + 0.5, (int)(uni(rng) * 40) - 12);
        const double t = (uni(rng) < 0.5 ? -1.0 : 1.0) * ldexp(uni(rng) + 0.5, (int)(uni(rng) * 30) - 20);
        check_row(s0, t, 300);
    }
    for (int c = 0; c < 400; c++) {  // long rows, through zero
        const double t = ldexp(uni(rng) + 0.5, (int)(uni(rng) * 8) - 4) * (uni(rng) < 0.5 ? -1.0 : 1.0);
        const double s0 = -t * (2000.0 + 6000.0 * uni(rng)) * (0.5 + uni(rng));
        check_sparse(s0, t, 9000, rng);
    }
    // ties at every addition: t = (k + 1/2) ulp of the start's binade, starts even and odd
    for (int c = 0; c < 2000; c++) {
        const int e = (int)(uni(rng) * 24) - 4;
        const double u = ldexp(1.0, e - 52);
        const double s0 = (uni(rng) < 0.5 ? -1.0 : 1.0) * (ldexp(1.0, e) + u * floor(uni(rng) * 4.0e15));
        const double k = floor(uni(rng) * 1.0e6) * (uni(rng) < 0.5 ? -1.0 : 1.0);
        check_row(s0, (k + 0.5) * u, 400);
    }
    // steps below half an ulp, exactly half an ulp, powers of two, zero
    for (int c = 0; c < 500; c++) {
        const double s0 = ldexp(uni(rng) + 1.0, (int)(uni(rng) * 20));
        const double u = ldexp(1.0, ilogb(s0) - 52);
        check_row(s0, u * 0.25, 50);
        check_row(s0, u * 0.5, 50);
        check_row(-s0, u * 0.5, 50);
        check_row(s0, ldexp(1.0, (int)(uni(rng) * 20) - 10), 200);
        check_row(s0, 0.0, 5);
    }
    // runs that reach either end of a binade, where the grid of the results changes
    for (int e = -6; e <= 30; e += 3) {
        const double u = ldexp(1.0, e - 52), top = ldexp(1.0, e + 1), bot = ldexp(1.0, e);
        for (int k = 1; k <= 24; k++)
            for (int j = 1; j <= 5; j++)
                for (double half = 0.0; half <= 0.5; half += 0.5)
                    for (double sg = -1.0; sg <= 1.0; sg += 2.0) {
                        check_row(sg * (top - k * u), sg * (j + half) * u, 60);
                        check_row(sg * (bot + k * u), -sg * (j + half) * u, 60);
                    }
    }
    check_row(0.0, 1.0, 100);
    check_row(-3.0, 1.0, 100);
    check_row(0.1, -0.1, 100);
    printf("seq_add: %ld values, %ld mismatches\n", g_cases, g_bad);
    return g_bad ? 1 : 0;
}
