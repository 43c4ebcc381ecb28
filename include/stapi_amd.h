/*
 * Extensions of the MI355X trace backend.  None of these has a counterpart
 * in the reference's coretrace/stapi.h; callers that only use the reference
 * interface never need them.
 */
#ifndef SOLTRACE_AMD_STAPI_AMD_H
#define SOLTRACE_AMD_STAPI_AMD_H

#include "stapi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Arithmetic of the closest-hit search: 64 (default, the reference's FP64
 * path) or 32 (FP32 search: BVH, element tests, Newton iterations and
 * aperture tests in float; frame transforms, sun-plane draws, every
 * accept/reject draw, the torus quartic, Newton start planes and the
 * per-hit finish -- hit point, normal, error rotations, reflection -- stay
 * FP64).  Returns 1, or -1 for other values. */
int st_amd_set_precision(st_context_t pcxt, int bits);

/* 1 (default): stage-1 sun rays use the sun-plane grid and other rays a
 * per-stage BVH; 0: every ray tests every element.  Both give the same hits
 * (ties break to the lower element index, as a full scan does). */
int st_amd_set_cull(st_context_t pcxt, int on);

/* Bin, during the next runs, the last record of each ray on element `elem`
 * (0-based) of stage `stage` (0-based) into an nx x ny map in that element's
 * frame, the way the GUI flux map does (app/src/project.cpp:1139-1436):
 * flat ('f') elements bin (x, y) over [xmin,xmax] x [ymin,ymax]; cylinders
 * ('t') and partial cylinders ('s' + 'l') unroll x to arc length, which fixes
 * the x range.  absorbed_only keeps only absorbed final records.  stage < 0
 * switches binning off.  Returns 1, or -1 for an element that cannot be
 * binned. */
int st_amd_flux_request(st_context_t pcxt, int stage, int elem, int nx, int ny, int absorbed_only, double xmin,
                        double xmax, double ymin, double ymax);

/* Integer counts of the last run's map, counts[ix*ny + iy]; bin_dims (may be
 * NULL) receives xmin, ymin, bin width, bin height.  Returns nx*ny or -1. */
int st_amd_flux_result(st_context_t pcxt, unsigned int *counts, double *bin_dims);

/* Per-source flux maps (on = 1) during the next runs: beside the map of
 * st_amd_flux_request, which must name an element of stage 2 or later, keep
 * one map per stage-1 element.  A ray counts in the map of its source, the
 * stage-1 element of its last stage-1 record (the reflection that sent it
 * on), so the maps add up to the requested map bin for bin.  The images a
 * flux surrogate is trained on: one receiver image per heliostat.  Rays
 * resumed from loaded stage data carry no source and count in no map.
 * Returns 1, or -1 for a bad context.  No counterpart in the reference. */
int st_amd_flux_by_source(st_context_t pcxt, int on);

/* Copy the last run's per-source maps into counts (capacity n words),
 * counts[(src * nx + ix) * ny + iy] for 0-based stage-1 element src.
 * counts may be NULL to ask for the size.  Returns the number of words
 * (stage-1 elements * nx * ny), or -1 when the last run kept none or n is
 * too small. */
long long st_amd_flux_source_result(st_context_t pcxt, unsigned int *counts, long long n);
/* The same maps copied device to device: counts is a DEVICE pointer on the
 * context's GPU, and the copy is queued on `stream` (a hipStream_t, NULL:
 * the null stream) without waiting.  The maps stay in HBM until the next
 * run.  Returns the number of words, or -1 as above. */
long long st_amd_flux_source_result_device(st_context_t pcxt, unsigned int *counts, long long n, void *stream);

/* The GUI's flux statistics of element `elem` (0-based) of stage `stage`
 * over the last run's records (app/src/project.cpp:1139-1436,
 * ElementStatistics::Compute and BinRaysXY), binned on the GPU where the
 * records are.  Flat ('f') and cylindrical ('t') surfaces and partial
 * cylinders ('s' with aperture 'l'), as the GUI allows.
 *   nx, ny         bins (> 1 each)
 *   autoscale      1: x/y range from every record on the element; 0: from
 *                  bounds
 *   absorbed_only  1: bin only absorbed last records
 *   dni            W/m2, for the power per ray
 *   bounds[4]      in (autoscale 0) and out: xmin, xmax, ymin, ymax, after
 *                  the cylinder unroll
 *   grid           (may be NULL) nx*ny ray counts, grid[ix*ny + iy]
 *   stats[15]      power per ray, peak flux, min flux, average flux, sigma,
 *                  uniformity, peak flux uncertainty %, average flux
 *                  uncertainty %, centroid x, y, z (element frame, to 1e-6 m),
 *                  rays binned, z value, bin size x, bin size y
 * Returns 1, or -1 for an element the GUI cannot bin, bad sizes, an empty
 * range or no run. */
int st_amd_element_stats(st_context_t pcxt, int stage, int elem, int nx, int ny, int autoscale, int absorbed_only,
                         double dni, double bounds[4], unsigned int *grid, double *stats);

/* st_amd_element_stats without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  The bins and accumulators live in a fixed scratch array on
 * each GPU, sized for at most 1,048,576 bins (nx*ny); larger grids return
 * -1 with a message.  grid is required (the counts are copied straight into
 * it).  Arguments and results are those of st_amd_element_stats.  Every call
 * on one GPU uses that GPU's one scratch array: calls that overlap in time
 * on the same GPU (from several threads) race and must be serialized by the
 * caller.  Returns 1 or -1. */
int st_amd_element_stats_noalloc(st_context_t pcxt, int stage, int elem, int nx, int ny, int autoscale,
                                 int absorbed_only, double dni, double bounds[4], unsigned int *grid,
                                 double *stats);

/* Trace n contexts' scenes together, context k with seed seeds[k], each with
 * the records, ray numbers and sun-ray count its own st_sim_run(ctxs[k],
 * seeds[k], NULL, NULL) gives (coretrace/stapi.cpp:680-734); the results
 * are read with the usual getters of each context.  No counterpart in the
 * reference, whose callers trace contexts one by one or on host threads
 * (app/src/trace.cpp:709-730).  Every kernel launch covers all the scenes:
 * persistent sun-mask and trace launches drain a device queue of per-wave
 * tiles drawn from every scene, so scenes of uneven size fill the GPU
 * together.  The trace launches' tile size and waves per SIMD are picked
 * per scene class by timing the candidates on the first launch of that
 * class and kept for the process.  All contexts must be on one device with
 * one precision, without flux requests or stage data; callbacks are not
 * called.  results[k] (may be NULL) receives each scene's record count or
 * -1 (messages on that context).  Returns 1, or -1 when any scene failed or
 * the arguments are invalid. */
int st_amd_sim_run_batch(st_context_t *ctxs, int n, const unsigned int *seeds, int *results);

/* Times and counts of the last batched run whose first context is pcxt:
 * [0] sun-mask ms, [1] sunshape ms, [2] wall ms, [3] sun rays, [4] mask
 * launches, [5] sunshape launches, [6..21] trace ms per stage, [22..37]
 * trace launches per stage, [38..53] rays per stage, [54..69] records per
 * stage.  Writes min(n, 70) values; returns 70, or -1 before any batch. */
int st_amd_batch_timing(st_context_t pcxt, double *out, int n);

/* The batch autotuner's choices so far, three ints each: class key
 * (precision bits * 100 + kernel variant * 10 + 1 for stages after the
 * first), rays per tile, waves per SIMD.  Writes min(n/3) entries; returns
 * the number of classes tuned. */
int st_amd_batch_tuning(int *out, int n);

/* Read a SolTrace .stinput system into the context with the headless strace
 * reader (coretrace/strace/main.cpp:358-397).  Returns 1, or -1 when the file
 * cannot be opened or is malformed. */
int st_amd_read_stinput(st_context_t pcxt, const char *path);

/* Write the last run's records to a file: format 0 is strace's .rays text
 * (coretrace/strace/main.cpp:399-483), format 1 the GUI's binary ray file
 * (app/src/project.cpp:915-948: six doubles SunXMin, SunXMax, SunYMin,
 * SunYMax, SunRayCount, Length, then nine doubles per record: x, y, z,
 * cos x, cos y, cos z, stage, element, ray number; native byte order).
 * Returns 1, or -1 on a bad format or a failed write. */
int st_amd_write_ray_file(st_context_t pcxt, const char *path, int format);

/* 0-based sun-ray index of rays 1..n of the last run (out[raynum-1]).
 * Returns the number written, or -1. */
int st_amd_ray_sun_index(st_context_t pcxt, unsigned long long *out, int n);

/* Device timing of the last run (HIP events): [0] sun-mask ms,
 * [1..16] trace ms per stage, [17] compaction ms, [18] wall ms of st_sim_run,
 * [19] sun rays drawn, [20..35] rays entering each stage, [36..51] records
 * per stage, [52] ms of the cooperative tabulated-sunshape draws.  Writes
 * min(n, 53) values; returns 53. */
int st_amd_timing(st_context_t pcxt, double *out, int n);

/* Count kernel events during the next runs (off by default: it adds
 * atomics).  Returns 1. */
int st_amd_count_events(st_context_t pcxt, int on);

/* Launch counts and event counters of the last run: [0] sun-mask launches,
 * [1..16] trace launches per stage, then five counters for the sun mask and
 * five per stage (rays, element tests, BVH nodes, uniforms drawn, records),
 * then [102] tries the tabulated-sunshape sampler evaluated and [103] the
 * tries the serial rejection loop would have made.  Writes min(n, 104)
 * values; returns 104. */
int st_amd_events(st_context_t pcxt, double *out, int n);

/* Stage-data save/resume: the reference's st_sim_run_data
 * (coretrace/stapi.cpp:680-728; raytrace.cpp:544-578, 1044-1078;
 * generateray.cpp:165-219) with plain arrays in place of its std::vector
 * arguments.
 *   s1, n1   rows of stage-1 records to load, 8 doubles each: x, y, z,
 *            cos x, cos y, cos z (stage frame), element, ray number
 *   s2, n2   rows of rays entering stage 2 to load, 7 doubles each: x, y, z,
 *            cos x, cos y, cos z (global frame), ray number
 * With n1 > 0 and n2 > 0, stage 1 is not traced. Its records are the rows
 * of s1, stage 2 starts from the rays of s2, and the sun-ray count is the
 * last s1 row's ray number. Random draws of a loaded ray use the stream of
 * sun-ray index (ray number - 1). n2 > 0 alone sets the ray count to n2, as
 * in the reference. A flux map requested on stage 1 (st_amd_flux_request)
 * bins the loaded records, each ray's last one on the element; a ray's rows
 * are taken to be contiguous, as saved.
 * With save_stage_data, the run keeps the rows the reference appends to its
 * two vectors; st_amd_saved_stage_data returns them.
 * Returns the record count, or -1 (messages as st_sim_run). */
int st_amd_sim_run_data(st_context_t pcxt, unsigned int seed, const double *s1, st_uint_t n1, const double *s2,
                        st_uint_t n2, int save_stage_data,
                        int (*callback)(st_uint_t ntracedtotal, st_uint_t ntraced, st_uint_t ntotrace,
                                        st_uint_t curstage, st_uint_t nstages, void *data),
                        void *data);

/* Rows saved by the last st_amd_sim_run_data with save_stage_data.
 *   which 1: every stage-1 record, in trace order (8 doubles a row, as s1);
 *   which 2: the reference's stage-2 ray array as stage 2 leaves it (7
 *            doubles a row, as s2). The reference reuses that array in
 *            place, so its first rows are the rays leaving stage 2 and the
 *            rest are the rays that entered stage 2 after them.
 * Writes the rows to out when it is not NULL; returns the row count, or -1
 * for another `which`. */
int st_amd_saved_stage_data(st_context_t pcxt, int which, double *out);

/* ---- Flux surrogate network (no counterpart in the reference) ----------
 * A multilayer perceptron, n_in -> depth x [width, ReLU] -> n_out, with
 * BF16 weights and activations, FP32 accumulation on the MFMA units, FP32
 * master weights and Adam, trained and evaluated by the library's own
 * kernels.  The flux surrogate maps a heliostat's pose, the sun and its aim
 * offset to its receiver flux image (the per-source maps above are its
 * training targets).  Data pointers (x, y, t) are DEVICE pointers on the
 * GPU current at creation, row-major, one row per sample; stream is a
 * hipStream_t (NULL: the null stream), and every call is ordered on it.
 * Calls return 1, or -1 with the reason in st_amd_mlp_error(). */

/* n_in <= 64; width and n_out multiples of 64; 1 <= depth <= 14; batches up
 * to max_batch.  Weights He-uniform from seed, biases zero.  NULL on error. */
void *st_amd_mlp_create(int n_in, int width, int depth, int n_out, int max_batch, unsigned long long seed);
void st_amd_mlp_free(void *net);
const char *st_amd_mlp_error(void);
long long st_amd_mlp_num_params(void *net);
/* Parameters in exchange order, host memory: per layer W [n_out][n_in]
 * row-major, then b [n_out].  which 0: FP32 master values; 1: the BF16
 * values the kernels multiply with.  set_params also clears Adam's state. */
int st_amd_mlp_get_params(void *net, float *out, int which);
int st_amd_mlp_set_params(void *net, const float *in);
/* y [batch][n_out] (f32) from x [batch][n_in] (f32, rounded to BF16). */
int st_amd_mlp_forward(void *net, const float *x, int batch, float *y, void *stream);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on the mean squared error
 * between the network's output for x and targets t [batch][n_out] (BF16
 * bits).  loss (host, may be NULL: no synchronisation) receives the loss
 * before the step. */
int st_amd_mlp_train_step(void *net, const float *x, const unsigned short *t, int batch, float lr, double *loss,
                          void *stream);
/* `steps` Adam steps on batches of a resident training set, replayed from
 * a HIP graph: one step (batch gather, forward, loss, backward, Adam) is
 * captured once per call and launched `steps` times on `stream`, reading
 * its batch and learning rate at a device step counter.  x [rows][n_in]
 * (f32) and t [rows][n_out] (BF16 bits) hold the set; idx [steps][batch]
 * (int32) the rows of each step's batch (clamped into [0, rows)); lr
 * [steps] (f32) each step's learning rate; all DEVICE pointers.  The steps
 * equal `steps` calls of st_amd_mlp_train_step on the same batches bit for
 * bit.  loss (host, may be NULL) receives the loss of the last step's batch
 * before its update.  Synchronises. */
int st_amd_mlp_fit(void *net, const float *x, const unsigned short *t, long long rows, const int *idx, int batch,
                   int steps, const float *lr, double *loss, void *stream);
/* Loss and its gradient with respect to every parameter (host, exchange
 * order) at the current weights, without a step.  Synchronises. */
int st_amd_mlp_gradients(void *net, const float *x, const unsigned short *t, int batch, float *grad, double *loss,
                         void *stream);

#ifdef __cplusplus
}

#include <vector>

/* The reference's own entry point, with the C++ linkage it has there
 * (coretrace/stapi.cpp:680-685; the declaration in stapi.h:159-162 is
 * commented out). Loads from, or appends saved rows to, the two vectors as
 * the reference does, by way of st_amd_sim_run_data. */
int st_sim_run_data(st_context_t pcxt, unsigned int seed, std::vector<std::vector<double> > *data_s1,
                    std::vector<std::vector<double> > *data_s2, bool save_stage_data,
                    int (*callback)(st_uint_t ntracedtotal, st_uint_t ntraced, st_uint_t ntotrace,
                                    st_uint_t curstage, st_uint_t nstages, void *data),
                    void *data);
#endif

#endif
