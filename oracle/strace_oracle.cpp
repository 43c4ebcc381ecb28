// Test infrastructure: the strace command line of soltrace_amd/strace run on
// the CPU oracle (st_oracle.so) instead of the HIP library.  With --counter
// first, the oracle draws from the per-ray counter streams and starts far
// torus solves where the HIP library does, so its records equal the HIP
// library's; without it, it draws from the reference's sequential stream and
// its .rays file equals the reference strace's.  With --binary it also
// writes the GUI's binary ray file to <system.stinput>.bin.  Both options come
// before the strace arguments, in either order.
#include <cstdio>
#include <cstring>
#include <string>

#include "stinput.h"

extern "C" int st_oracle_set_rng(void *cxt, int mode);
extern "C" int st_oracle_set_torus_start(void *cxt, int on);

static void counter_mode(st_context_t cxt) {
    st_oracle_set_rng(cxt, 1);
    st_oracle_set_torus_start(cxt, 1);
}

static void write_binary(st_context_t cxt, const char *file) {
    const std::string out = std::string(file) + ".bin";
    if (!st_io::write_rays_binary(out.c_str(), cxt)) printf("failed to write %s\n", out.c_str());
}

int main(int argc, char **argv) {
    void (*prepare)(st_context_t) = nullptr;
    void (*after)(st_context_t, const char *) = nullptr;
    int k = 1;
    for (; k < argc; k++) {
        if (strcmp(argv[k], "--counter") == 0) prepare = counter_mode;
        else if (strcmp(argv[k], "--binary") == 0) after = write_binary;
        else break;
    }
    argv[k - 1] = argv[0];
    return st_io::strace_main(argc - (k - 1), argv + (k - 1), prepare, after);
}
