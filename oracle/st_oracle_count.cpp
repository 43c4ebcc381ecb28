// The oracle with operation counting (st_count.h).  TEST AND MEASUREMENT
// INFRASTRUCTURE ONLY: bench.py runs it on a bounded sample of each workload
// to give the algorithmic floating-point work per ray of the trace loop, the
// figure the GPU kernels' achieved rates are priced against.  Same C API as
// st_oracle.so (every double there is a Counted here, which passes through
// the ABI as a double does), plus st_opcount_reset / st_opcount_read.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <strings.h>
#include <unordered_map>
#include <utility>
#include <vector>

#include "st_count.h"

#define double opcount::Counted
#include "st_oracle.cpp"
#undef double

extern "C" {
// Zero every phase's counts.
void st_opcount_reset() {
    opcount::State &s = opcount::state();
    memset(s.n, 0, sizeof s.n);
    memset(s.tests, 0, sizeof s.tests);
    s.phase = s.saved = 2;
}
// counts[phase * 5 + column], columns add/sub, mul, div, sqrt,
// transcendental, then the element tests of stages 1, 2 and 3+; writes
// min(n, 83) values, returns 83.
int st_opcount_read(unsigned long long *out, int n) {
    const opcount::State &s = opcount::state();
    int m = 0;
    for (int p = 0; p < opcount::kPhases; p++)
        for (int c = 0; c < opcount::kCols; c++)
            if (m < n) out[m++] = s.n[p][c];
    for (int k = 0; k < 3; k++)
        if (m < n) out[m++] = s.tests[k];
    return opcount::kPhases * opcount::kCols + 3;
}
}
