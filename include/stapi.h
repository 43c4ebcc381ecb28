/*
 * C API of the MI355X trace backend: the entry points of SolTrace's
 * coretrace/stapi.h, with identical names, argument types and return
 * conventions, exported by soltrace_amd/coretrace_api.so (and the static
 * libcoretrace.a) so that callers written against the reference's header
 * (pysoltrace's ctypes layer, the GUI, strace) link or load it unchanged.
 *
 * Each declaration cites the reference interface it replaces.  Return
 * conventions (coretrace/stapi.cpp): NULL context -> -1 (NULL for
 * st_message); bad stage/element/optic index -> -1; success -> 1; the
 * st_add_* calls return the new index; st_sim_run returns the record count
 * or -1; result getters return -2 when a copy fails.
 *
 * Differences a caller can observe:
 *  - st_sim_run traces on the GPU.  Random draws come from per-ray
 *    counter-based streams (Philox4x32-10 keyed by seed, stage and ray), so
 *    records match the reference statistically, not draw-for-draw; hit
 *    geometry given the same draws matches the reference's FP64 path.
 *  - records stay in GPU memory until a getter is called.
 */
#ifndef SOLTRACE_AMD_STAPI_H
#define SOLTRACE_AMD_STAPI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef unsigned long st_uint_t;  /* coretrace/stapi.h:72 (64-bit on LP64) */
typedef void *st_context_t;       /* coretrace/stapi.h:73 */

/* contexts: coretrace/stapi.h:76-77 */
st_context_t st_create_context();
int st_free_context(st_context_t pcxt);

/* messages: coretrace/stapi.h:80-82 */
int st_num_messages(st_context_t pcxt);
const char *st_message(st_context_t pcxt, st_uint_t idx);
void st_clear_messages(st_context_t pcxt);

/* optics: coretrace/stapi.h:92-105 */
int st_num_optics(st_context_t pcxt);
int st_add_optic(st_context_t pcxt, const char *name);
int st_delete_optic(st_context_t pcxt, st_uint_t idx);
int st_clear_optics(st_context_t pcxt);
int st_optic(st_context_t pcxt, st_uint_t idx, int fb, char dist, int optnum, int apgr, int order, double rreal,
             double rimag, double ref, double tra, double gratingab12[3], double rmsslope, double rmsspec,
             int userefltable, int refl_npoints, double *refl_angles, double *refls, int usetranstable,
             int trans_npoints, double *trans_angles, double *transs);

/* stages: coretrace/stapi.h:108-116 */
int st_num_stages(st_context_t pcxt);
int st_add_stage(st_context_t pcxt);
int st_add_stages(st_context_t pcxt, st_uint_t num);
int st_delete_stage(st_context_t pcxt, st_uint_t idx);
int st_clear_stages(st_context_t pcxt);
int st_stage_xyz(st_context_t pcxt, st_uint_t idx, double x, double y, double z);
int st_stage_aim(st_context_t pcxt, st_uint_t idx, double ax, double ay, double az);
int st_stage_zrot(st_context_t pcxt, st_uint_t idx, double zrot);
int st_stage_flags(st_context_t pcxt, st_uint_t idx, int virt, int multihit, int tracethrough);

/* elements: coretrace/stapi.h:119-136 */
int st_num_elements(st_context_t pcxt, st_uint_t stage);
int st_add_element(st_context_t pcxt, st_uint_t stage);
int st_add_elements(st_context_t pcxt, st_uint_t stage, st_uint_t num);
int st_delete_element(st_context_t pcxt, st_uint_t stage, st_uint_t idx);
int st_clear_elements(st_context_t pcxt, st_uint_t stage);
int st_element_enabled(st_context_t pcxt, st_uint_t stage, st_uint_t idx, int enabled);
int st_element_xyz(st_context_t pcxt, st_uint_t stage, st_uint_t idx, double x, double y, double z);
int st_element_aim(st_context_t pcxt, st_uint_t stage, st_uint_t idx, double ax, double ay, double az);
int st_element_zrot(st_context_t pcxt, st_uint_t stage, st_uint_t idx, double zrot);
int st_element_aperture(st_context_t pcxt, st_uint_t stage, st_uint_t idx, char ap);
int st_element_aperture_params(st_context_t pcxt, st_uint_t stage, st_uint_t idx, double params[8]);
int st_element_surface(st_context_t pcxt, st_uint_t stage, st_uint_t idx, char surf);
int st_element_surface_params(st_context_t pcxt, st_uint_t stage, st_uint_t idx, double params[8]);
int st_element_surface_file(st_context_t pcxt, st_uint_t stage, st_uint_t idx, const char *file);
int st_element_interaction(st_context_t pcxt, st_uint_t stage, st_uint_t idx, int type);
int st_element_optic(st_context_t pcxt, st_uint_t stage, st_uint_t idx, const char *name);

/* sun: coretrace/stapi.h:139-142 */
int st_sun(st_context_t pcxt, int point_source, char shape, double sigma_halfwidth);
int st_sun_xyz(st_context_t pcxt, double x, double y, double z);
int st_sun_position(st_context_t pcxt, double lat, double day, double hour, double *x, double *y, double *z);
int st_sun_userdata(st_context_t pcxt, st_uint_t npoints, double angle[], double intensity[]);

/* results: coretrace/stapi.h:145-151 */
int st_num_intersections(st_context_t pcxt);
int st_locations(st_context_t pcxt, double *loc_x, double *loc_y, double *loc_z);
int st_cosines(st_context_t pcxt, double *cos_x, double *cos_y, double *cos_z);
int st_elementmap(st_context_t pcxt, int *element_map);
int st_stagemap(st_context_t pcxt, int *stage_map);
int st_raynumbers(st_context_t pcxt, int *ray_numbers);
int st_sun_stats(st_context_t pcxt, double *xmin, double *xmax, double *ymin, double *ymax, int *nsunrays);

/* simulation: coretrace/stapi.h:154-157 */
int st_sim_params(st_context_t pcxt, int raycount, int maxcount, int include_dynamic_group);
int st_sim_errors(st_context_t pcxt, int include_sun_shape, int include_optics);
int st_sim_run(st_context_t pcxt, unsigned int seed,
               int (*callback)(st_uint_t ntracedtotal, st_uint_t ntraced, st_uint_t ntotrace, st_uint_t curstage,
                               st_uint_t nstages, void *data),
               void *data);

/* utilities: coretrace/stapi.h:166-171 */
void st_calc_euler_angles(double origin[3], double aimpoint[3], double zrot, double euler[3]);
void st_transform_to_local(double posref[3], double cosref[3], double origin[3], double rreftoloc[3][3],
                           double posloc[3], double cosloc[3]);
void st_transform_to_reference(double posloc[3], double cosloc[3], double origin[3], double rloctoref[3][3],
                               double posref[3], double cosref[3]);
void st_matrix_vector_mult(double m[3][3], double v[3], double mxv[3]);
void st_calc_transform_matrices(double euler[3], double rreftoloc[3][3], double rloctoref[3][3]);
void st_matrix_transpose(double input[3][3], double output[3][3]);

#ifdef __cplusplus
}
#endif

#endif
