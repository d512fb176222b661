/*
 * find_zig_ties.c -- writes tests/golden/zig_ties.json: stream positions where
 * a ziggurat attempt draw u ties the fast test on its high word, i.e.
 * (u >> 32) == (T[l] >> 21) for its layer l = u & 0xff.  On such a draw the
 * GPU samplers' 32-bit fast test cannot decide and the exact 53-bit test
 * (normal.rs:26-29: (u >> 11) * 2^-53 * X[l] < X[l+1], equivalently
 * (u >> 11) < T[l]) settles it; one draw in 2^32 ties, so the tests need
 * these positions found.  Test infrastructure: it reuses the oracle's
 * stream and ziggurat (oracle/stocha_oracle.c, included).
 *
 *   gcc -O2 -fopenmp -ffp-contract=off -I oracle -o /tmp/find_zig_ties \
 *       tests/golden/find_zig_ties.c -lm && /tmp/find_zig_ties > /tmp/zig_ties.json
 *
 * (threads print in any order; the committed fixture lists the ties sorted by
 * path.)  About 13 s on 8 cores.
 *
 * Searches seed 42, block 4096, paths 0 .. 2^22-1, the first 4000 normals of
 * each path, and records every tie at an attempt draw (draw index within the
 * path, normal index, and whether the exact test accepts).
 */
#include "../../oracle/stocha_oracle.c"

static const uint64_t ZT[256] = {ZIG_T_VALUES};

#define SEED 42ull
#define BLOCK 4096ull
#define PATHS (1ull << 22)
#define NORMALS 4000

int main(void) {
    printf("{\"seed\": %llu, \"block\": %llu, \"paths_searched\": %llu, \"normals_per_path\": %d,\n"
           " \"ties\": [\n", (unsigned long long)SEED, (unsigned long long)BLOCK,
           (unsigned long long)PATHS, NORMALS);
    int first = 1;
#pragma omp parallel for schedule(dynamic, 256)
    for (long long p = 0; p < (long long)PATHS; p++) {
        pcg_t g;
        pcg_for_path(&g, SEED, BLOCK, (uint64_t)p);
        uint64_t k = 0; /* draws taken */
        for (int j = 0; j < NORMALS; j++) {
            for (;;) { /* normal.rs:18-51 */
                uint64_t u = pcg_next_u64(&g);
                uint64_t at = k++;
                int l = (int)(u & 0xFF);
                uint64_t m = u >> 11;
                if ((u >> 32) == (ZT[l] >> 21)) {
#pragma omp critical
                    {
                        printf("%s  {\"path\": %lld, \"draw\": %llu, \"normal\": %d, \"accepted\": %s}",
                               first ? "" : ",\n", p, (unsigned long long)at, j,
                               m < ZT[l] ? "true" : "false");
                        first = 0;
                        fflush(stdout);
                    }
                }
                if (m < ZT[l]) break;
                if (l == 0) { /* normal.rs:54-67 */
                    for (;;) {
                        double u1 = pcg_next_f64(&g), u2 = pcg_next_f64(&g);
                        k += 2;
                        if (u1 <= 0.0 || u2 <= 0.0) continue;
                        double x = -log(u1) / ZR;
                        if (-2.0 * log(u2) >= x * x) break;
                    }
                    break;
                }
                double x = (double)m * (1.0 / 9007199254740992.0) * ZX[l];
                double y = ZY[l] + pcg_next_f64(&g) * (ZY[l + 1] - ZY[l]);
                k++;
                if (y < exp(-0.5 * x * x)) break;
            }
        }
    }
    printf("\n]}\n");
    return 0;
}
