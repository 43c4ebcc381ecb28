// SolTrace .stinput reader and .rays writer over the public C API.
//
// The reader follows the headless strace reader (coretrace/strace/main.cpp:
// read_sun :76-128, read_optic_surface :130-230, read_optic :233-249,
// read_element :252-314, read_stage :317-356, read_system :358-397) and the
// writer its write_data_file (:399-483).  Both call nothing but the stapi.h
// functions, so they drive any library that exports that API.
#pragma once
#include <cstdio>

#include "stapi.h"

namespace st_io {

// Read a whole .stinput system into the context.  Progress lines go to log
// (may be null).  False on a malformed file or a setter that fails.
bool read_system(FILE *fp, st_context_t cxt, FILE *log);

// Write the context's last run as a .rays text file.  False when the file
// cannot be opened.
bool write_rays(const char *path, st_context_t cxt);

// Write the context's last run as the GUI's binary ray file
// (app/src/project.cpp:915-948, RayData::WriteDataFile): six doubles
// SunXMin, SunXMax, SunYMin, SunYMax, SunRayCount, Length, then nine doubles
// per record: x, y, z, cos x, cos y, cos z, stage, element, ray number.
// False when the file cannot be written.
bool write_rays_binary(const char *path, st_context_t cxt);

// The strace command line: <system.stinput> <rays> <max rays> <seed>
// <sunshape 0/1> <errors 0/1> [<power tower 0/1>].  Reads the system, runs
// it and writes <system.stinput>.rays.  Returns the process exit code.
// prepare (may be null) is called on the context after the system is read
// and before the run; after (may be null) once the run succeeded, with the
// system's path.
int strace_main(int argc, char **argv, void (*prepare)(st_context_t cxt),
                void (*after)(st_context_t cxt, const char *file) = nullptr);

}  // namespace st_io
