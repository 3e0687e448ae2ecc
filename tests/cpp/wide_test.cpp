// Host build of the GPU exact stage for general doubles (exact_wide.h +
// ExactTT<WidePolicy>, the code k_exact_wide runs), driven by the pytest in
// tests/test_wide_exact.py.
// stdin: one pair per line, 18 doubles in C99 hex-float form (triangle A's
// three vertices, then B's). stdout, per pair:
//   "R"                                  frame or endpoints do not fit (host residual)
//   "P hit coplanar degenerate nseg exp2 shifted npred" then nseg lines
//   "S n0 .. n7"                         each number as 4 hex words, low first
#include <stdio.h>

#include "exact_wide.h"

using namespace gi;

static void print_w(const W256& x) {
    for (int i = 0; i < 4; ++i) printf(" %016llx", (unsigned long long)x.w[i]);
}

int main() {
    double v[18];
    while (true) {
        int got = 0;
        for (int i = 0; i < 18; ++i) got += scanf("%la", &v[i]);
        if (got != 18) break;
        WideExact::V3 A[3], B[3];
        WideFrame f;
        if (!wide_frame(v, v + 9, A, B, f)) {
            printf("R\n");
            continue;
        }
        // the fast pass first, as on the GPU; coplanar pairs take the full one
        WideExact::Fast fast;
        WideExact::intersect_fast(A, B, fast);
        WideExact::Result r;
        WideExact::intersect(A, B, r);
        if (fast.deferred != r.coplanar || (!fast.deferred && fast.hit != r.hit)) {
            printf("X fast/full disagree\n");
            return 1;
        }
        WideExact::HP P[GI_SEG_MAX][2];
        bool fits = true;
        for (int s = 0; s < r.nseg; ++s) {
            if (!r.coplanar) {
                P[s][0] = fast.p;
                P[s][1] = fast.q;
            } else {
                WideExact::segment(A, B, r, s, P[s][0], P[s][1]);
            }
            fits = fits && wide_absolute(f, P[s][0]) && wide_absolute(f, P[s][1]);
        }
        if (!fits) {
            printf("R\n");
            continue;
        }
        printf("P %d %d %d %d %d %d %d\n", r.hit, r.coplanar, r.degenerate, r.nseg, f.exp2, f.shifted, r.npred);
        for (int s = 0; s < r.nseg; ++s) {
            printf("S");
            for (int e = 0; e < 2; ++e) {
                for (int k = 0; k < 3; ++k) print_w(P[s][e].h[k]);
                print_w(P[s][e].w);
            }
            printf("\n");
        }
    }
    return 0;
}
