// C entry points of the .stinput reader and the ray-file writers, built into
// coretrace_api.so beside the C API they call.
#include "stinput.h"

#define ST_EXPORT extern "C" __attribute__((visibility("default")))

ST_EXPORT int st_amd_read_stinput(st_context_t pcxt, const char *path) {
    if (!pcxt || !path) return -1;
    FILE *fp = fopen(path, "r");
    if (!fp) return -1;
    const bool ok = st_io::read_system(fp, pcxt, nullptr);
    fclose(fp);
    return ok ? 1 : -1;
}

ST_EXPORT int st_amd_write_ray_file(st_context_t pcxt, const char *path, int format) {
    if (!pcxt || !path) return -1;
    if (format == 0) return st_io::write_rays(path, pcxt) ? 1 : -1;
    if (format == 1) return st_io::write_rays_binary(path, pcxt) ? 1 : -1;
    return -1;
}
