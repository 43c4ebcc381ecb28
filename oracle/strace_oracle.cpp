// Test infrastructure: the strace command line of soltrace_amd/strace run on
// the CPU oracle (st_oracle.so) instead of the HIP library.  With --counter
// first, the oracle draws from the per-ray counter streams and starts far
// torus solves where the HIP library does, so its records equal the HIP
// library's; without it, it draws from the reference's sequential stream and
// its .rays file equals the reference strace's.
#include <cstring>

#include "stinput.h"

extern "C" int st_oracle_set_rng(void *cxt, int mode);
extern "C" int st_oracle_set_torus_start(void *cxt, int on);

static void counter_mode(st_context_t cxt) {
    st_oracle_set_rng(cxt, 1);
    st_oracle_set_torus_start(cxt, 1);
}

int main(int argc, char **argv) {
    if (argc > 1 && strcmp(argv[1], "--counter") == 0) {
        argv[1] = argv[0];
        return st_io::strace_main(argc - 1, argv + 1, counter_mode);
    }
    return st_io::strace_main(argc, argv, nullptr);
}
