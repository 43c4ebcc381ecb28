// strace: trace a SolTrace .stinput system headlessly on the MI355X backend
// and write <system.stinput>.rays (the command line of
// coretrace/strace/main.cpp:501-561).
//
//   strace <system.stinput> <rays> <max rays> <seed> <sunshape 0/1> <errors 0/1> [<power tower 0/1>]
#include "stinput.h"

int main(int argc, char **argv) { return st_io::strace_main(argc, argv, nullptr); }
