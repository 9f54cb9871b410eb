/*
 * gls.h -- C ABI of libgls.so, the MI355X (gfx950) generalized Lomb-Scargle
 * periodogram via the adjoint NFFT.
 *
 * Plain extern "C" functions, plain pointers and sizes.  Every call returns an
 * int status (GLS_OK == 0, negative GLS_E_* on failure); gls_last_error()
 * gives a thread-local description of the last failure.  No C++ exception
 * crosses this boundary.  Device pointers are plain addresses of device
 * memory (from gls_malloc or any HIP allocation on the same device); streams
 * are hipStream_t values passed as void*.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * johnh2o2/cuvarbase 0.3.0).  The reference binds its kernels through PyCUDA
 * `prepare`d functions and cuFFT through scikit-cuda; this library replaces
 * that whole launch sequence per curve (precompute_psi -> fast_gaussian_grid
 * -> nfft_shift -> cuFFT inverse -> normalize, twice, then lomb) by a planned,
 * batched pipeline over many curves.
 */
#ifndef CUVARBASE_AMD_GLS_H
#define CUVARBASE_AMD_GLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GLS_OK              0
#define GLS_E_INVALID      -1   /* bad argument (size, pointer, mode ...)     */
#define GLS_E_HIP          -2   /* HIP runtime error                          */
#define GLS_E_OOM          -3   /* device or host allocation failed           */
#define GLS_E_NOT_INIT     -4   /* gls_init() not called                      */
#define GLS_E_UNSUPPORTED  -5   /* size or option outside what is built       */
#define GLS_E_DATA         -6   /* input data violates a documented precondition */

/* ---- periodogram modes (cuvarbase/kernels/lomb.cu:13-15) --------------- */
#define GLS_MODE_STANDARD       0   /* lspow0: no offset term                */
#define GLS_MODE_FLOATING_MEAN  1   /* lspow_flmean: GLS with offset (default) */
#define GLS_MODE_WINDOW         2   /* spectral window                       */

/* ---- input kinds -------------------------------------------------------- */
/* RAW: t (f64), y (f64), dy (f64).  The device normalises and weights each
 * curve as normalize_light_curves + LombScargleMemory.setdata do
 * (cuvarbase/utils.py:113-146, cuvarbase/lombscargle.py:235-293).          */
#define GLS_INPUT_RAW       0
/* WEIGHTED: t (f64), yw (f32), w (f32) already prepared on the host, with
 * ybar and yy given per curve -- the argument set of the reference's `lomb`
 * and `lomb_dirsum` kernels (cuvarbase/kernels/lomb.cu:144-147, :189-197).  */
#define GLS_INPUT_WEIGHTED  1

/* ---- methods ------------------------------------------------------------ */
#define GLS_WINDOW_GAUSSIAN 0
#define GLS_WINDOW_ES       1   /* exponential of semicircle: same accuracy at a smaller
                                   sigma and m in FP32; no reference counterpart */

#define GLS_METHOD_NFFT     0   /* use_fft=True  (cuvarbase/lombscargle.py:660-684) */
#define GLS_METHOD_DIRECT   1   /* use_fft=False (cuvarbase/lombscargle.py:643-659) */
/* direct sums at explicit, arbitrarily spaced frequencies: the reference's
 * lomb_dirsum_custom_frq kernel (cuvarbase/kernels/lomb.cu:167-187), which
 * its Python layer never launches; run with gls_lomb_execute_custom */
#define GLS_METHOD_DIRECT_CUSTOM 2

/* One light curve of a batch.  Observations of curve i are
 * [obs_offset, obs_offset + n_obs) of the input arrays (any order); its nf
 * powers go to power[out_offset .. out_offset+nf).  Output ranges of
 * different curves must not overlap.
 * The frequency grid is f_k = f0 + k*df, k = 0..nf-1, evaluated exactly
 * (no rounding of f0 to a multiple of df). */
typedef struct gls_curve {
    int64_t obs_offset;
    int32_t n_obs;
    int32_t nf;
    double  f0;
    double  df;
    int64_t out_offset;
    double  ybar;      /* GLS_INPUT_WEIGHTED only: weighted mean of y       */
    double  yy;        /* GLS_INPUT_WEIGHTED only: weighted variance of y   */
} gls_curve;

typedef struct gls_params {
    double  sigma;       /* NFFT oversampling; grid n >= sigma*nf (cuvarbase/cunfft.py:352) */
    int32_t m;           /* Gaussian half-width in grid cells (cuvarbase/cunfft.py:353)     */
    int32_t mode;        /* GLS_MODE_*                                                      */
    double  reg_c;       /* amplitude-prior lambda on the cosine amplitude, reg[0]          */
    double  reg_s;       /* amplitude-prior lambda on the sine amplitude,   reg[1]          */
                         /* (cuvarbase/lombscargle.py:103-118)                              */
    int32_t input_kind;  /* GLS_INPUT_*                                                     */
    int32_t method;      /* GLS_METHOD_*                                                    */
    int32_t window;      /* GLS_WINDOW_GAUSSIAN (the reference's, cuvarbase/cunfft.py:56-57)
                            or GLS_WINDOW_ES (exp(beta (sqrt(1 - z^2) - 1)) on 2m+1 cells) */
    float   window_beta; /* ES shape; <= 0: pi (2m+1)(1 - 1/(2 sigma))                  */
    int64_t grid_chunk_bytes; /* > 0: the NFFT grids of a size bucket run in chunks of at
                           most this many bytes that reuse one workspace region (workspace
                           shrinks from all grids of the batch to the largest chunk, at a
                           few per cent of throughput); <= 0: one region per curve.
                           No reference counterpart. */
    int64_t point_capacity; /* observations the plan's per-point workspace holds; <= the
                           batch's own total: exactly that total.  Larger capacities let
                           gls_plan_rebind put other curves (more points) on the plan.
                           No reference counterpart. */
    int32_t precision;   /* GLS_PRECISION_*; GLS_PRECISION_FP32 needs method NFFT */
    int32_t spread;      /* GLS_SPREAD_*: how pass 1 accumulates the window taps */
} gls_params;

/* ---- spreading accumulation (gls_params.spread) ------------------------- */
/* FP32 atomic adds in LDS (the default): the order of additions depends on
 * scheduling, so results can differ in the last bits from run to run and
 * between a curve alone and in a batch (as the reference's global float
 * atomics, cuvarbase/kernels/cunfft.cu:153-161) */
#define GLS_SPREAD_FLOAT 0
/* fixed point: every tap is rounded to an integer multiple of 2^-e and added
 * with LDS integer atomics, which are exact and associative.  All the cells a
 * point touches in a pass-1 tile lie in one grid row, so e is chosen per tile
 * row from the number of points in the row times their largest |value|
 * (no cell can exceed 2^30): accuracy as FP32 accumulation, and results
 * bitwise the same for a curve alone, in any batch and on every run.  Costs a
 * counting pass over each tile's points and a conversion of its cells: about
 * 18% more pass-1 time on the 5,000-curve survey batch. */
#define GLS_SPREAD_FIXED 1
/* per curve: fixed point for curves of at least one point per 80 grid cells
 * on pass-1 columns of 64 cells or more, whose dense tiles the integer adds
 * fill faster than FP32 atomics (pass 1 1.3-3x faster there), FP32 atomics
 * for the others (fixed point costs 10-30% more on sparse tiles, and more on
 * short columns, where the per-row counting and conversion weigh most).  Each curve's choice depends on its own point count and
 * grid only (gls_plan_curve_spread reports it; gls_plan_rebind keeps each
 * curve slot's mode); fixed-point curves are bitwise reproducible, the others
 * as GLS_SPREAD_FLOAT.  A size bucket with curves of both kinds runs two
 * pass-1 launches and one pass-2 launch.  No reference counterpart. */
#define GLS_SPREAD_AUTO 2

/* ---- precision of the time coordinates (gls_params.precision) ----------- */
/* FP64 time offsets, grid cells and modulation phases, FP64 low-frequency
 * fix-up (the default) */
#define GLS_PRECISION_FP64_COORDS 0
/* the reference's arithmetic for comparison: t rounded to FP32 after centring
 * on its mean (cuvarbase/lombscargle.py:259), time offsets, grid coordinates
 * and modulation phases in FP32 (cuvarbase/kernels/cunfft.cu:101-103,
 * :233-241), no FP64 low-frequency fix-up */
#define GLS_PRECISION_FP32        1

/* One adjoint-NFFT item: sum_j y_j exp(2 pi i (f0 + k df) t_j), k < nf. */
typedef struct gls_nfft_item {
    int64_t obs_offset;
    int32_t n_obs;
    int32_t nf;
    double  f0;
    double  df;
    int64_t out_offset;  /* offset in complex elements */
} gls_nfft_item;

/* One curve of a multi-harmonic power evaluation (gls_mh_power). */
typedef struct gls_mh_curve {
    int64_t sums_offset; /* first complex element of the curve's 3H sum rows   */
    int64_t out_offset;  /* first output element                                */
    int32_t nf;          /* frequencies K                                       */
    int32_t pad;
    double  yy;          /* weighted variance YY                                */
    double  ybar;        /* sum of yw (the reference subtracts ybar*C, ybar*S)   */
    double  sum_w;       /* sum of the weights, c[0] (1 for normalised weights)  */
    double  lam;         /* amplitude-prior lambda = prior^-2 on every harmonic  */
} gls_mh_curve;

typedef struct gls_plan gls_plan;   /* opaque */

/* ---- context ------------------------------------------------------------ */
/* Replaces the import-time context creation `import pycuda.autoprimaryctx`
 * (cuvarbase/__init__.py:2): selects `device`.  Idempotent per device. */
int gls_init(int device);
int gls_shutdown(void);
const char* gls_last_error(void);
int gls_device_count(int* count);
/* Reports name (<= len bytes), number of CUs and total memory in bytes. */
int gls_device_info(char* name, int len, int* num_cus, size_t* total_mem);
int gls_version(int* major, int* minor, int* patch);

/* ---- streams and events: cuda.Stream / synchronize (cuvarbase/core.py:27-40) */
int gls_stream_create(void** stream);
int gls_stream_destroy(void* stream);
int gls_stream_synchronize(void* stream);
/* *done = 1 if all work on the stream has completed, else 0. */
int gls_stream_query(void* stream, int* done);
int gls_event_create(void** event);
int gls_event_destroy(void* event);
int gls_event_record(void* event, void* stream);
int gls_event_synchronize(void* event);
int gls_event_query(void* event, int* done);
int gls_event_elapsed_ms(void* start, void* stop, float* ms);
int gls_device_synchronize(void);

/* ---- memory: gpuarray.zeros / set_async / get_async / aligned_zeros
 *      (cuvarbase/lombscargle.py:129-191, :295-308) ----------------------- */
int gls_malloc(void** ptr, size_t bytes);
int gls_free(void* ptr);
int gls_host_alloc(void** ptr, size_t bytes);   /* page-locked host memory */
int gls_host_free(void* ptr);
int gls_memcpy_h2d_async(void* dst, const void* src, size_t bytes, void* stream);
int gls_memcpy_d2h_async(void* dst, const void* src, size_t bytes, void* stream);
int gls_memcpy_d2d_async(void* dst, const void* src, size_t bytes, void* stream);
/* Copy by a kernel between device memory and page-locked host memory from
 * gls_host_alloc (either end may be either), enqueued on `stream` like a
 * launch.  No counterpart in the reference (whose copies are PyCUDA's
 * set_async/get_async, cuvarbase/lombscargle.py:295-308): for small,
 * latency-critical copies, since the DMA path (gls_memcpy_h2d_async) has
 * been measured holding the calling thread for 7-9 ms now and then in a
 * served stream of small batches.  GLS_E_INVALID for memory the device
 * cannot reach (pageable host memory). */
int gls_memcpy_kernel_async(void* dst, const void* src, size_t bytes, void* stream);
int gls_memset_async(void* dst, int value, size_t bytes, void* stream);
int gls_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ---- GLS periodogram ---------------------------------------------------- */
/* Plans a batch: chooses each curve's grid size n >= sigma*nf of the form
 * n1 * 2^p2 (n1 = 2^a 3^b 5^c from a fixed catalogue, 2^10 <= n <= 2^24)
 * and lays out the workspace.  Replaces LombScargleMemory.allocate +
 * NFFTMemory.allocate_grid + the two cuFFT plans per curve
 * (cuvarbase/lombscargle.py:129-181, cuvarbase/cunfft.py:82-92).
 * `curves` is read during the call only.  At most 2^22 curves per plan
 * (GLS_E_UNSUPPORTED above: split the batch), so that no per-curve launch
 * exceeds 2^31 work-items. */
int gls_lomb_plan_create(gls_plan** plan, const gls_curve* curves,
                         int32_t ncurves, const gls_params* params);
/* Plans are host objects (no device memory): their descriptors are copied
 * into the first bytes of the workspace by every execute.  Destroying a plan
 * after its execute calls have returned is safe. */
int gls_plan_destroy(gls_plan* plan);
/* Device workspace the plan needs (caller allocates, e.g. gls_malloc). */
int gls_plan_workspace_bytes(const gls_plan* plan, size_t* bytes);
/* Total number of output elements (sum of nf) the plan writes. */
int gls_plan_output_count(const gls_plan* plan, int64_t* count);
/* Grid size chosen for curve i (0 for the direct method). */
int gls_plan_grid_size(const gls_plan* plan, int32_t i, int64_t* n);
/* Four-step split n = n1 * n2 and Gaussian width b (cells^2) of curve i
 * (zeros for the direct method). */
int gls_plan_curve_geometry(const gls_plan* plan, int32_t i, int64_t* n,
                            int32_t* n1, int32_t* n2, double* b);
/* Tiles the plan uses for curve i's size bucket: c1 grid columns per pass-1
 * workgroup, cr grid rows per pass-2 workgroup (zeros for the direct method). */
int gls_plan_curve_tiles(const gls_plan* plan, int32_t i, int32_t* c1, int32_t* cr);
/* How pass 1 accumulates curve i's window taps: GLS_SPREAD_FLOAT or
 * GLS_SPREAD_FIXED (the plan's choice under GLS_SPREAD_AUTO; FLOAT for the
 * direct method).  No reference counterpart. */
int gls_plan_curve_spread(const gls_plan* plan, int32_t i, int32_t* spread);

/* ---- tiles of the NFFT passes (no counterpart in the reference, whose FFTs
 * are cuFFT plans and whose gridding kernel has a fixed block size) --------
 * A size bucket is a four-step split (n1, p2), n = n1 * 2^p2.  Plans take a
 * bucket's tiles from a table measured on the device and compiled into the
 * library (default rules for buckets it does not list).  gls_tile_set
 * overrides them for plans created afterwards, process-wide: c1 = 0 or
 * cr = 0 keeps the library's choice for that pass, both 0 remove the
 * override.  GLS_E_UNSUPPORTED if the library has no kernel for the tile
 * (a library built for tile sweeps has one for every candidate: c1 and cr
 * powers of two up to 16, n1 * c1 <= 8192, 256 <= 2^p2 * cr <= 4096).
 * gls_tile_get returns the tiles a plan made now would use. */
int gls_tile_set(int32_t n1, int32_t p2, int32_t c1, int32_t cr);
int gls_tile_get(int32_t n1, int32_t p2, int32_t* c1, int32_t* cr);

/* Work-items of one launch of an NFFT pass: at most 2^31, the default (a
 * dispatch counts its work-items in 32 bits).  A size bucket with more runs as
 * several launches over consecutive curves; a smaller cap (set before the
 * execute, for every plan) splits launches further, as tests of that path
 * do.  items <= 0 restores the default.  No reference counterpart. */
int gls_launch_items_set(int64_t items);
/* A plan gives each curve the grid n = n1 * 2^p2 >= sigma * nf of least
 * n x (cost per cell of its bucket): the time of the two passes per cell
 * measured on the device (the same table), or a model (one pass over the
 * cells plus 0.12 per FFT radix stage, scaled to the table) for buckets it
 * does not list.  gls_grid_cost_set overrides a bucket's cost in
 * picoseconds per cell for plans created afterwards (<= 0 removes the
 * override); gls_grid_cost_get returns the cost in use and its source
 * (0 model, 1 measured table, 2 override). */
int gls_grid_cost_set(int32_t n1, int32_t p2, double ps_per_cell);
int gls_grid_cost_get(int32_t n1, int32_t p2, double* ps_per_cell, int32_t* source);

/* Runs the batch on `stream`: replaces lomb_scargle_async for every curve
 * (cuvarbase/lombscargle.py:579-689: nfft_adjoint_async twice + `lomb`, or
 * `lomb_dirsum` for the direct method).
 *   t      device f64 observation times
 *   a, b   RAW: y (f64), dy (f64);  WEIGHTED: yw (f32), w (f32)
 *   power  device f32 output, sum(nf) values; invalid powers are -1
 *          (cuvarbase/kernels/lomb.cu:79-80)
 *   workspace  device buffer of gls_plan_workspace_bytes() bytes
 * Asynchronous; returns after enqueueing. */
int gls_lomb_execute(gls_plan* plan, const double* t, const void* a,
                     const void* b, float* power, void* workspace,
                     void* stream);

/* gls_lomb_execute without a workspace argument and without any allocation:
 * no reference counterpart.  For callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The workspace is a static device array of the library, one
 * per GPU, of GLS_STATIC_SCRATCH_BYTES (128 MiB), held from the library's
 * loading to the end of the process; plans whose gls_plan_workspace_bytes is
 * larger fail with GLS_E_UNSUPPORTED.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (different streams, or threads) race
 * and must be serialised by the caller; a call may begin once the previous
 * call's work on the device has been ordered before it (same stream, or an
 * event wait).  Nothing in the array needs clearing between calls: every
 * execute initialises what it reads.  Size buckets run in series on `stream`
 * (gls_plan_set_fork is ignored).  Calls no hipMalloc/hipFree; the
 * descriptors go up through the library's page-locked staging, allocated by
 * gls_init. */
#define GLS_STATIC_SCRATCH_BYTES (128u << 20)
int gls_lomb_execute_static(gls_plan* plan, const double* t, const void* a,
                            const void* b, float* power, void* stream);

/* Run the independent size buckets of an NFFT plan concurrently on the
 * library's auxiliary streams (forked from and joined back into the execute
 * stream): a few per cent faster on mixed batches; kernel durations then
 * overlap.  Off by default.  No reference counterpart. */
int gls_plan_set_fork(gls_plan* plan, int32_t enable);

/* Captured executes: the kernel sequence of gls_lomb_execute(_f64) on fixed
 * pointers as a HIP graph, for launch-bound loops (the reference launches
 * nine kernels per curve from Python, cuvarbase/lombscargle.py:579-689).
 * Creation copies the plan's descriptors into `workspace` (once) and
 * captures the launches on `stream` (a created stream); every
 * gls_graph_launch then recomputes the powers from the current contents of
 * t, a, b.  The workspace belongs to the graph until it is destroyed.  The
 * plan may be destroyed after creation.  Lomb-Scargle plans of methods NFFT
 * and DIRECT. */
typedef struct gls_graph gls_graph;   /* opaque */
int gls_plan_graph_create(gls_plan* plan, const double* t, const void* a,
                          const void* b, void* power, int32_t power_is_f64,
                          void* workspace, void* stream, gls_graph** graph);
int gls_graph_launch(gls_graph* graph, void* stream);
int gls_graph_destroy(gls_graph* graph);

/* Repeated batch shapes (no reference counterpart; the reference allocates and
 * plans every curve anew, cuvarbase/lombscargle.py:1065-1074).
 * gls_plan_rebind replaces the curves of a Lomb-Scargle plan by `curves`:
 * the same count, each with the frequency grid (nf, f0, df) of the curve it
 * replaces, on the same side of the 16,384-point preparation limit, no more
 * points in all than the plan's point_capacity (and wide curves no longer
 * than the plan's longest); observation and output offsets, sizes and the
 * weighted-input ybar/yy may change.  Host-side only: executes afterwards use
 * the new curves.  gls_graph_update enqueues, on `stream`, the upload of the
 * plan's descriptors into the workspace of a graph captured from that plan,
 * so the next gls_graph_launch on the stream computes the rebound curves
 * (from the graph's fixed input and output pointers). */
int gls_plan_rebind(gls_plan* plan, const gls_curve* curves, int32_t ncurves);
int gls_graph_update(gls_graph* graph, const gls_plan* plan, void* stream);

/* Custom-frequency plans (GLS_METHOD_DIRECT_CUSTOM): curve i is evaluated
 * at freqs[out_offset .. out_offset + nf) (device f64; f0/df ignored) into
 * f32 (power_is_f64 = 0) or f64 powers.  Replaces lomb_dirsum_custom_frq
 * (cuvarbase/kernels/lomb.cu:167-187). */
int gls_lomb_execute_custom(gls_plan* plan, const double* t, const void* a,
                            const void* b, const double* freqs, void* power,
                            int32_t power_is_f64, void* workspace, void* stream);

/* Same, into an f64 power array (use_double=True,
 * cuvarbase/lombscargle.py:96-101). */
int gls_lomb_execute_f64(gls_plan* plan, const double* t, const void* a,
                         const void* b, double* power, void* workspace,
                         void* stream);

/* ---- per-stage timing (no counterpart in the reference, whose only timing
 *      is wall clock around whole calls, docs/source/plots/benchmarks.py:53-63).
 * When enabled, gls_lomb_execute records HIP events on its stream between
 * the launches of the plan's stages: prep, then per grid-size bucket pass 1
 * and pass 2 (or the direct-sum launch).  After the stream has completed,
 * gls_plan_stage_times reports for each stage its kind (0 prep, 1 pass 1,
 * 2 pass 2, 3 direct sums), grid size n of its bucket, elapsed ms and the
 * algorithmic bytes it must move. */
int gls_plan_enable_timing(gls_plan* plan, int32_t enable);
int gls_plan_stage_count(const gls_plan* plan, int32_t* count);
int gls_plan_stage_times(const gls_plan* plan, int32_t* kind, int32_t* grid,
                         float* ms, double* bytes, int32_t max);
/* The size bucket a pass-1 or pass-2 stage runs: column length n1, log2 of
 * the row length p2, tiles c1 and cr, GLS_SPREAD_FLOAT or GLS_SPREAD_FIXED
 * (its pass-1 kernel's accumulation) and streamed = 1 when pass 1 stores the
 * bucket's grids non-temporally (grids of more than 96 MB, which the 256 MB
 * memory-side cache would not keep until pass 2); zeros and GLS_SPREAD_FLOAT
 * for the other stages.  No reference counterpart. */
int gls_plan_stage_bucket(const gls_plan* plan, int32_t stage, int32_t* n1, int32_t* p2,
                          int32_t* c1, int32_t* cr, int32_t* spread, int32_t* streamed);

/* ---- adjoint NFFT (NFFTAsyncProcess.run, cuvarbase/cunfft.py:507-542) --- */
/* y is f64 (or f32 with y_is_f32 = 1); output is complex64 (interleaved
 * re, im) at out_offset*2 floats per item. */
int gls_nfft_plan_create(gls_plan** plan, const gls_nfft_item* items,
                         int32_t nitems, double sigma, int32_t m);
int gls_nfft_execute(gls_plan* plan, const double* t, const void* y,
                     int32_t y_is_f32, float* out, void* workspace,
                     void* stream);

/* Gaussian gridding only, in the reference's one-sided geometry: real grid of
 * ng = int(sigma*nf) cells, x = (t - tmin)/(spp (tmax - tmin)); the
 * `just_return_gridded_data=True` output of nfft_adjoint_async
 * (cuvarbase/cunfft.py:229-268).  grid must hold ng floats and is
 * overwritten.  n_obs points from t (f64) and y (f64). */
int gls_gaussian_grid(const double* t, const double* y, int32_t n_obs,
                      int32_t ng, int32_t m, double b, double tmin,
                      double tmax, double spp, float* grid, void* stream);

/* ---- multi-harmonic GLS (next row: nharmonics > 1) ------------------------ */
/* Multi-harmonic floating-mean GLS power at K frequencies per curve from the
 * harmonic sums that an NFFT plan computed: for curve i, `sums` holds 3H rows
 * of K complex values starting at curves[i].sums_offset:
 *   rows 0 .. H-1:   sum yw exp(2 pi i n f_k t),  n = 1 .. H
 *   rows H .. 3H-1:  sum w  exp(2 pi i n f_k t),  n = 1 .. 2H.
 * Builds CC, CS, SS, YC, YS as mhdirect_sums (cuvarbase/lombscargle.py:337-400),
 * adds the amplitude prior (add_regularization, :403-447) and returns
 * P = (2 theta.b - theta^T A theta)/YY (mhgls_from_sums, :497-536); the
 * reference's LombScargleAsyncProcess rejects nharmonics > 1
 * (cuvarbase/lombscargle.py:727-730) and its lomb_mh kernel is a stub
 * (cuvarbase/kernels/lomb.cu:223-254).  `curves` is a DEVICE array;
 * 1 <= nharm <= 4; power f32 (power_is_f64 = 0) or f64; invalid -> -1. */
int gls_mh_power(const void* sums, const gls_mh_curve* curves, int32_t ncurves,
                 int32_t nharm, int32_t max_nf, void* power, int32_t power_is_f64,
                 void* stream);

/* ---- peak extraction (next row: best frequencies, lombscargle.py:1099-1107) */
/* For each curve of a plan executed into `power` with `workspace`: the top-k
 * powers and their indices, ignoring indices where mask[out_offset + k] != 0
 * (mask may be NULL).  top_power: ncurves*k f32, top_index: ncurves*k i32
 * (descending power, ties by lower index; -1 where fewer than k valid
 * values).  Enqueue on the stream of the execute (or after it). */
int gls_topk(const gls_plan* plan, const void* workspace, const float* power,
             const uint8_t* mask, int32_t k, float* top_power,
             int32_t* top_index, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* CUVARBASE_AMD_GLS_H */
