/*
 * gpucpm.h -- C ABI of the MI355X Cellular Potts engine (libgpucpm.so).
 *
 * Each entry point replaces one method of the reference's host engine
 * (shabaz/gpu-cpm src/cpm.h:78-112, src/cpm.cu) as reached through its
 * CPython binding (src/python_wrapper.cpp).  Plain C types only: no torch,
 * no HIP types.  Every call returns 0 on success or a negative CPM_E* code;
 * cpm_last_error() then holds a message (per thread).  One engine per
 * handle; a handle is not thread-safe.  The engine owns all memory;
 * outputs go to caller buffers.
 */
#ifndef GPUCPM_H
#define GPUCPM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPM_OK 0
#define CPM_EINVAL -1   /* bad argument (shape, type, schedule, ...) */
#define CPM_EHIP -2     /* HIP runtime error (no device, out of memory, ...) */
#define CPM_ESTATE -3   /* call out of order (e.g. run before push) */
#define CPM_ENOTIMPL -4 /* accepted by the reference, not implemented here */

#define CPM_NTYPES 20   /* settings.h:4 numTypes */

typedef struct cpm_engine* cpm_handle;
typedef struct cpm_batch* cpm_batch_handle;

/* Keyword arguments of Cpm.run (python_wrapper.cpp:186-216, cpm.cu:291-296).
 * Only positions_per_checkerboard, updates_per_checkerboard_switch,
 * iterations and inner_iterations change results; the rest select a thread
 * mapping in the reference and are validated but do not change results
 * here.  The RNG seed is set separately (cpm_set_seed). */
typedef struct {
    int cell_sync;
    int block_sync;
    int global_sync;
    int threads_per_block;
    int positions_per_thread;
    int positions_per_checkerboard;
    int updates_per_checkerboard_switch;
    int updates_per_barrier;
    int iterations;
    int inner_iterations;
    int shared;
    int partial_dispatch;
    int centroid_tracking;
} cpm_run_args;

/* Message of the last failed call on this thread ("" if none). */
const char* cpm_last_error(void);

/* Library build string (for diagnostics). */
const char* cpm_version(void);

/* Cpm::Cpm (cpm.cu:176-261), argument order of PyCpm_init
 * (python_wrapper.cpp:27-41): dimension L (power of two), dimensionality d
 * (2 or 3), number_of_cells n (tables hold n+1 rows, row 0 = medium),
 * temperature, has_field, history, has_act. */
int cpm_create(int dimension, int dimensionality, int number_of_cells, int temperature,
               int has_field, int history, int has_act, cpm_handle* out);

/* Cpm::~Cpm (cpm.cu:263-285). */
int cpm_destroy(cpm_handle h);

/* Cpm::setTemperature (cpm.cu:57-60); real temperatures are an extension. */
int cpm_set_temperature(cpm_handle h, double temperature);

/* Cpm::setArea (cpm.cu:125-129): area (volume in 3D) target and lambda. */
int cpm_set_area(cpm_handle h, int type, int target, double lambda);

/* Cpm::setPerimeter (cpm.cu:119-123): perimeter (surface in 3D). */
int cpm_set_perimeter(cpm_handle h, int type, int target, double lambda);

/* Cpm::setAdhesion (cpm.cu:137-143); J is stored as int16 (settings.h:20). */
int cpm_set_adhesion(cpm_handle h, int type, int other_type, int adhesion, int set_both);

/* Extension: adhesion as a real number; it must be a multiple of 2^-8 (the
 * energy is kept in dyadic fixed point).  Used by parameter fitting. */
int cpm_set_adhesion_real(cpm_handle h, int type, int other_type, double adhesion, int set_both);

/* Cpm::setStickiness (cpm.cu:145-148); int16 (settings.h:22). */
int cpm_set_stickiness(cpm_handle h, int type, int other_type, int stickiness);

/* Cpm::setChemotaxis (cpm.cu:155-156): lambda of the chemotaxis term
 * -lambda (d . f) (kernels2d.cu:325-334, kernels3d.cu:314-325), d the unit
 * vector of the copy, f the field at the source site.  |lambda| <= 8192. */
int cpm_set_chemotaxis(cpm_handle h, int type, double lambda);
/* Cpm::setAct (cpm.cu:131-135): lambda and actMax of the act term
 * -lambda/actMax (G_source - G_target) (kernels2d.cu:312-343,386-397,
 * 417-473,532-534), G the geometric mean of site activities over a site and
 * its same-cell neighbours.  max_act in [0, 65535].  Runs with the term need
 * has_act.  With shared = 1 and positions_per_checkerboard = 1 (the term
 * then reads sites other attempts of the phase write) each phase is decided
 * against its start lattice, then applied: two launches per phase (with
 * cell_sync the decide launch also claims, the apply launch checks the
 * claims); in a batch only in the per-phase form (else CPM_ENOTIMPL). */
int cpm_set_act(cpm_handle h, int type, int max_act, double lambda);
/* Cpm::setPersistence (cpm.cu:112-115): the persistence term -lambda (d . p)
 * (kernels3d.cu:327-344, 3D only as in the reference), p the cell's preferred
 * direction, updated after every inner_iterations chunk from the centroid
 * history of `history` chunks with the type's diffusion (kernels.cu:158-212,
 * cpm.cu:487-514).  time is stored (the reference's device path ignores it).
 * |lambda| <= 8192. */
int cpm_set_persistence(cpm_handle h, int type, int time, double diffusion, double lambda);
/* Extension: preferred directions, (n+1)*3 doubles (row 0 unused).  They
 * start as unit vectors drawn from the seed (cpm_set_seed re-draws them). */
int cpm_get_pref_dirs(cpm_handle h, double* out);
int cpm_set_pref_dirs(cpm_handle h, const double* in);
/* Cpm::getActData (cpm.cu:522-525): host act lattice, L^d values (the MCS of
 * the last copy into each site); pushed and pulled with the lattice.
 * CPM_ESTATE if the engine was created with has_act = 0. */
int cpm_host_act(cpm_handle h, uint32_t** out);
/* Cpm::getField (cpm.cu:527-528): host field, d * L^d doubles, axis-major
 * ([axis][site]); pushed with the lattice, quantized to 2^-16 on the device
 * (|f| < 2^15).  CPM_ESTATE if the engine was created with has_field = 0. */
int cpm_host_field(cpm_handle h, double** out);
/* Cpm::setFixed (cpm.cu:150-153): a fixed type is never overwritten, and a
 * copy from a fixed source writes medium instead (kernels2d.cu:281-290). */
int cpm_set_fixed(cpm_handle h, int type, int fixed);

/* Cpm::addCell (cpm.cu:160-174): writes id (k | type<<24) for the k-th
 * added cell into the host lattice.  z < 0 selects the 2D form. */
int cpm_add_cell(cpm_handle h, int type, int x, int y, int z);

/* Cpm::getCellIds (cpm.cu:519-521): the engine-owned host lattice,
 * L^d uint32 in [z][y][x] order, valid until cpm_destroy. */
int cpm_host_lattice(cpm_handle h, uint32_t** out);

/* Number of cells registered with cpm_add_cell (Cpm::_activeCells). */
int cpm_active_cells(cpm_handle h, int* out);

/* Cpm::moveCellIdsToGpu (cpm.cu:541-597): copy the host lattice to the
 * device and rebuild every cell table from it. */
int cpm_push(cpm_handle h);

/* Cpm::run (cpm.cu:291-517): (iterations / inner_iterations) *
 * inner_iterations Monte Carlo steps.  Asynchronous (like the 2D reference
 * run): returns once the work is queued. */
int cpm_run(cpm_handle h, const cpm_run_args* args);

/* Cpm::synchronize (cpm.cu:653-655). */
int cpm_sync(cpm_handle h);

/* Cpm::retreiveCellIdsFromGpu (cpm.cu:531-540): device lattice -> host. */
int cpm_pull(cpm_handle h);

/* Cpm::get2dCentroids / get3dCentroids (cpm.cu:68-110): rows for cells
 * 1..active_cells, d doubles each ([x, y(, z)]).  capacity is the row
 * capacity of out; *rows receives active_cells. */
int cpm_get_centroids(cpm_handle h, double* out, int capacity, int* rows);

/* Extension (the reference keeps these tables device-only): raw integer
 * cell tables for all n+1 rows.  area/perim: [n+1]; anchor: [(n+1)*3]
 * coordinates of each cell's first site in raster order at the last push;
 * csum: [(n+1)*3] sums of site coordinates unwrapped about the anchor.
 * Any pointer may be NULL. */
int cpm_get_cell_table(cpm_handle h, int32_t* area, int32_t* perim, int32_t* anchor,
                       int64_t* csum);

/* Extension: the random stream.  mode 0 (default): the counter stream, Philox
 * keyed by (sub-block, MCS, phase, attempt); with positions_per_checkerboard
 * = 1 two x-neighbouring same-colour sub-blocks share one Philox block.  mode 1: the reference's stream
 * (random.cu:5-28 seeded per thread at kernels2d.cu:99 / kernels3d.cu:86),
 * for comparisons with the reference: one MINSTD state per reference
 * thread of positions_per_thread sites per axis, seeded at every
 * inner_iterations chunk from the thread's first site index + the low 32
 * bits of the seed, consumed in the reference's draw order; acceptance
 * u < ceil(exp(-E/T) (2^31 - 1)).  With shared = 1 and positions_per_thread
 * > 2c the reference's layout is kept (kernels2d.cu:72-86,92-106,136-146):
 * threads in blocks of threads_per_block per axis, each thread's
 * positions_per_thread / 2c regions of 2c sites per axis visited one block
 * slice at a time; threads_per_block x positions_per_thread must divide
 * dimension (CPM_EINVAL otherwise).
 * A batch's engines must share the mode (CPM_EINVAL otherwise); the
 * persistent batch form does not run it (CPM_ENOTIMPL). */
/* Extension: online autotuning of the per-phase kernel form (no counterpart
 * in the reference, whose kernel is fixed at build time).  Every form
 * computes the same results, so while a workload (dimension, size, schedule,
 * types, terms) is unsettled each launch takes the candidate form with the
 * fewest samples, timed with events; after 8 samples of each the fastest is
 * kept for later runs of that workload (the choices are shared by the
 * engines of one process).  Off by default. */
int cpm_set_autotune(cpm_handle h, int on);
/* Attempt counters, the reference's analytics block (cpm.cu:215,559: three
 * device counters allocated and zeroed; their increments at
 * kernels2d.cu:292,526,531 and kernels3d.cu:263,548,553 are commented out)
 * made live behind a flag.  cpm_set_counters(h, 1) allocates and zeroes
 * them (0 switches counting off; the counts stay readable).
 * cpm_get_counters fills out[5]: attempts drawn (L^d per MCS since the
 * counters were switched on or reset), attempts whose source and target
 * differ, attempts evaluated (after the fixed-type rules; analytics[0]),
 * proposals accepted (analytics[1]), flips committed (after cell_sync's
 * claim check; analytics[2]); reset != 0 zeroes them.  Each attempt counts
 * once, also in claimed or split phases; batched runs count per engine.
 * Results never depend on the flag. */
int cpm_set_counters(cpm_handle h, int on);
int cpm_get_counters(cpm_handle h, int64_t out[5], int reset);
int cpm_set_rng(cpm_handle h, int mode);
int cpm_get_rng(cpm_handle h, int* mode);

/* Extension: RNG seed (Philox4x32-10 key) and the MCS counter that keys the
 * stream; together they make every run reproducible and resumable. */
int cpm_set_seed(cpm_handle h, uint64_t seed);
int cpm_get_seed(cpm_handle h, uint64_t* seed);
int cpm_set_mcs(cpm_handle h, uint64_t mcs);
int cpm_get_mcs(cpm_handle h, uint64_t* mcs);

/* Extension (no counterpart in the reference): kernel timing.  Between
 * begin and end every sweep launch is bracketed by HIP events recorded at
 * the dispatch's start and end on the engine's stream; end synchronizes and
 * returns the summed kernel time, the launch count and the number of copy
 * attempts (site-updates) those launches made. */
int cpm_timing_begin(cpm_handle h);
int cpm_timing_end(cpm_handle h, double* kernel_ms, int64_t* launches, int64_t* site_updates);

/* Extension (no counterpart in the reference): launch form of the last run.
 * form: 0 = one launch per colour phase, 1 = one cooperative launch per
 * inner_iterations chunk with the lattice in global memory, 2 = the same
 * with 2D tiles resident in LDS; blocks = cooperative grid size; tile = LDS
 * tile edge (0 if none).  Results never depend on the form. */
int cpm_last_launch(cpm_handle h, int* form, int* blocks, int* tile);
/* Extension: name of the sweep kernel the last run launched, as profilers
 * show it (e.g. "k_sweep_phase_compact<2, true, false>"); "" before any run. */
const char* cpm_last_kernel(cpm_handle h);

/* Extension: choose the launch form (-1 = automatic, else a form number as
 * above, 3 = one workgroup with the lattice in LDS, 4 = resident 2D tiles
 * that wait only on their neighbours between colour phases) and the LDS tile
 * edge (0 = automatic; form 4 takes 32 or 64) for this engine's runs; the
 * autotuner sets these.  Results never depend on them. */
int cpm_set_launch(cpm_handle h, int form, int tile);
/* Extension: chunks of resident-form runs (launch form 4) that were redone
 * with one launch per phase: *guard counts those where a cell spread over
 * more than two tiles per axis (the form's exactness condition), *waits
 * those where a neighbour wait gave up (either pointer may be NULL). */
int cpm_resident_fallbacks(cpm_handle h, int64_t* guard, int64_t* waits);
/* Extension (no counterpart in the reference, whose kernel variant is chosen
 * by template flags, cpm.cu:13-45): options of the per-phase kernels for
 * this engine's runs, each -1 (default), 0 or 1: small = read at most four
 * types' parameters with scalar loads (default in 3D); compact = with u = 1
 * queue only the attempts whose sites differ (default 1); full = the queue
 * carries whole neighbourhoods (default 0); attempts = attempts drawn per
 * thread by the compacted kernel, -1 or 1, 2, 4, 8 (default 2 in 2D, 8 in
 * 3D); narrow = 16-bit sites when they fit (default 1); box3 = 3D with c = 1
 * stages 32-bit boxes in LDS (default 0).  Results never depend on them. */
int cpm_set_kernel_options(cpm_handle h, int small, int compact, int full, int attempts,
                           int narrow, int box3);

/* Extension (no counterpart in the reference, whose parameters are one
 * global __constant__ block, kernels.cu:48-53): batched replicas.  A batch
 * is a fixed list of pushed engines of one dimensionality, each with its own
 * lattice size, parameters, temperature, seed and MCS counter.
 * cpm_batch_run applies one set of run arguments to all of them: by
 * default with one launch per colour phase covering every replica (form 0),
 * or with one cooperative launch per inner_iterations chunk in which each
 * replica's blocks form a group with its own barrier (form 1, 2D only).
 * Every replica ends bit-identical to running cpm_run on it alone.  The
 * engines must outlive the batch; their streams are ordered after the
 * batch's work. */
int cpm_batch_create(const cpm_handle* sims, int n, cpm_batch_handle* out);
int cpm_batch_run(cpm_batch_handle b, const cpm_run_args* args);
int cpm_batch_sync(cpm_batch_handle b);
/* Persistent form: blocks per unit of work (one block-width of
 * sub-blocks) given to each replica's group, default 1; the autotuner sets
 * it.  Results never depend on it. */
int cpm_batch_set_blocks(cpm_batch_handle b, double blocks_per_work);
/* Extension: the autotuner of cpm_set_autotune for batched runs (attempts
 * drawn per thread 1 / 4 / 8, staged or scalar parameters). */
int cpm_batch_set_autotune(cpm_batch_handle b, int on);
/* Batch launch form: 0 = one launch per colour phase, 1 = persistent.
 * Extension: 2 = the per-phase launches of up to 64 MCS captured once as a
 * hip graph (a device MCS counter advanced inside it) and replayed; the graph
 * is kept while runs keep the same shape; not with cell_sync, the reference
 * stream, persistence or kernel timing, where form 0 runs instead.  -1 =
 * choose: form 0, or with cpm_batch_set_autotune whole runs timed per MCS
 * across forms 0, 2 and (where it applies) 1, the fastest kept.  Results
 * never depend on the form. */
int cpm_batch_set_form(cpm_batch_handle b, int form);
/* Extension: the form the last batched run used (0, 1 or 2), and the graph
 * captures and graph replays of this batch so far. */
int cpm_batch_graph_info(cpm_batch_handle b, int* last_form, int64_t* captures, int64_t* replays);
/* Extension: cpm_set_kernel_options for batched runs (small, compact and
 * attempts per thread: -1 or 1, 4, 8; default 4 in 2D, 8 in 3D).  Each
 * replica's own narrow option still applies.  Results never depend on them. */
int cpm_batch_set_kernel_options(cpm_batch_handle b, int small, int compact, int attempts);
int cpm_batch_info(cpm_batch_handle b, int* replicas, int* blocks);
/* Extension: kernel timing of a batch's per-phase launches, as
 * cpm_timing_begin / cpm_timing_end for one engine (site_updates counts the
 * copy attempts of every replica). */
int cpm_batch_timing_begin(cpm_batch_handle b);
int cpm_batch_timing_end(cpm_batch_handle b, double* kernel_ms, int64_t* launches, int64_t* site_updates);
/* Extension: name of the per-phase batch kernel of the last batched run ("" before one). */
const char* cpm_batch_last_kernel(cpm_batch_handle b);
int cpm_batch_destroy(cpm_batch_handle b);

/* Extension (no counterpart in the reference): sufficient statistics of
 *   H(sigma; theta) = sum_{unordered neighbour pairs, s_i != s_j} J[t_i][t_j]
 *                   + sum_{cells c>0} lam_A[t_c] (A_c - A0[t_c])^2
 *                   + sum_{cells c>0} lam_P[t_c] (P_c - P0[t_c])^2
 * on the current device lattice, as exact integers that do not depend on
 * theta: pairs[a*20 + b] (a <= b) counts unlike-cell neighbour pairs of
 * types {a, b}; moments[5*t + k] holds, over live cells of type t,
 * k = 0: count, 1: sum A, 2: sum A^2, 3: sum P, 4: sum P^2.  H and dH/dtheta
 * for every theta follow in closed form (the backward pass of the energy). */
int cpm_energy_stats(cpm_handle h, int64_t* pairs, int64_t* moments);
/* Extension: cpm_energy_stats of n engines at once (one statistics launch
 * for all of them, after each engine's queued work), into pairs[n * 400]
 * and moments[n * 100]. */
int cpm_energy_stats_many(const cpm_handle* sims, int n, int64_t* pairs, int64_t* moments);
/* Extension with no counterpart in the reference, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): cpm_energy_stats_many with no
 * allocation.  Its statistics land in a file-scope device scratch array of
 * 64 engines x 500 words (256 KiB, held by each GPU until the process ends),
 * so it takes at most 64 engines (CPM_EINVAL beyond); it runs on the first
 * engine's stream after every engine's queued work, clears the scratch with
 * a fill kernel and waits for the result.  Calls that overlap on one GPU
 * share the scratch and race: serialize them. */
int cpm_energy_stats_many_noalloc(const cpm_handle* sims, int n, int64_t* pairs, int64_t* moments);
/* Extension: kernel time (ms, HIP events on the statistics stream) of the
 * last cpm_energy_stats / cpm_energy_stats_many call: one launch per
 * dimensionality and site width (16- or 32-bit) covers every engine of the
 * call. */
int cpm_energy_stats_timing(double* kernel_ms);
/* Extension (no counterpart in the reference): the backward pass in one call,
 * H(sigma; theta) and dH/dtheta on the engine's current lattice for theta
 * laid out, for ntypes types, as: J upper triangle (a <= b, row-major) |
 * lambda_area[t] | target_area[t] | lambda_perimeter[t] | target_perimeter[t]
 * (nt (nt + 1) / 2 + 4 nt doubles; dtheta alike).  H is the adhesion, area
 * and perimeter energy (stickiness enters dH only); the derivatives are
 * exact polynomials in theta over integer statistics, in double. */
int cpm_energy_grad(cpm_handle h, int ntypes, const double* theta, double* H, double* dtheta);

#ifdef __cplusplus
}
#endif
#endif
