// Internal declarations shared by the libgls.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gls.h"
#include "sizes.hpp"
#include "tiles.hpp"

namespace gls {

// ---------------------------------------------------------------------------
// error handling (thread-local message, integer status at the C ABI)
// ---------------------------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define GLS_HIP_TRY(expr)                                                     \
    do {                                                                      \
        hipError_t _e = (expr);                                               \
        if (_e != hipSuccess) return ::gls::hip_fail(_e, #expr);              \
    } while (0)

bool initialised();
int require_init();

// Asynchronous host-to-device copy of a plan's descriptors through the
// library's page-locked staging ring (allocated by gls_init): the source is
// copied into the ring on the host and the DMA reads the ring, so the call
// never waits on the runtime's shared pageable staging buffer.  Falls back to
// a plain hipMemcpyAsync for blocks larger than half the ring.
hipError_t upload_async(void* dst, const void* src, size_t bytes, hipStream_t s);

// Device address and size of the static scratch array of the current device
// (gls_lomb_execute_static); GLS_E_NOT_INIT before gls_init.
constexpr size_t kScratchBytes = size_t(128) << 20;
int static_scratch(void** ptr);

// Device copy of the ES deconvolution table for (m, beta), uploaded once per
// process and kept until it ends (plans point at it; it is not part of the
// per-execute descriptors).
int es_table_device(int m, float beta, const float** out);

// library-owned streams (default priority) onto which an execute forks the
// independent size buckets of a plan; created on first use
constexpr int kAuxStreams = 3;
hipError_t aux_streams(hipStream_t* out);

constexpr int kMaxGrid = 1 << 24;   // largest grid n (cells)
constexpr int kMinGrid = 1 << 10;   // smallest grid n

// ---------------------------------------------------------------------------
// Inter-pass grid layout (pass 1 writes, pass 2 reads): the rows k1 of a
// grid in blocks of kRowBlock, each block stored row-block-major over the n2
// columns, A[(k1 / R) R n2 + u2 R + k1 % R] (R = kRowBlock).  A pass-1 column
// is then written as R x 8-byte segments and the rows of a pass-2 tile lie in
// one contiguous run per row block (DRAM pages are opened once per tile,
// not once per column).  A grid takes ceil(n1 / R) R n2 cells.  R = 8 (the
// pass-2 tile height of the survey's main buckets) measured 4% faster in pass
// 2 than R = 16 on the survey batch, with pass 1 unchanged.
#ifndef GLS_ROW_BLOCK
#define GLS_ROW_BLOCK 8
#endif
constexpr int kRowBlock = GLS_ROW_BLOCK;
static_assert((kRowBlock & (kRowBlock - 1)) == 0, "the row block is a power of two");
__host__ __device__ inline int64_t grid_cells(int n1, int p2) {
    return (int64_t)((n1 + kRowBlock - 1) / kRowBlock * kRowBlock) << p2;
}

// per-curve plan data (host computed, uploaded once per plan)
// ---------------------------------------------------------------------------
struct CurveGeom {
    int64_t obs;       // first observation index in t / a / b
    int64_t out;       // first output element
    int64_t pts;       // first point slot in the per-point workspace arrays
    int64_t bkt;       // first int of this curve's bucket table (2*(nb+1))
    int64_t grid;      // first float2 of this curve's grids (ngrids * grid_cells)
    double f0;
    double df;
    double b;          // Gaussian width in cells^2
    double ybar_in;    // GLS_INPUT_WEIGHTED
    double yy_in;      // GLS_INPUT_WEIGHTED
    int32_t n_obs;
    int32_t nf;        // K, number of output frequencies
    int32_t n;         // grid size n = n1 * n2
    int32_t n1;        // pass-1 column length
    int32_t p2;        // log2 of pass-2 row length n2
    int32_t c1;        // pass-1 columns per tile
    int32_t kc;        // K/2, band centre index
    int32_t ci;        // index of the curve in the plan
    // per-curve constants of the pass kernels, computed on the host (a
    // division or square root in every thread would cost tens of VALU
    // instructions per block)
    float g_binv;      // log2(e) / b        (Gaussian spreading weight)
    float g_norm;      // 1 / sqrt(pi b)
    float g_bpi2;      // b pi^2             (Gaussian deconvolution)
    float ninv;        // 1 / n (FP32)
};

// written by the prep kernel
struct CurveState {
    double tref;       // time origin (min t)
    double ybar;
    double yy;
    double tspan;      // max t - min t
    double tmean;      // mean of the finite t (FP32-coordinate plans only)
};

// per-point workspace, indexed by CurveGeom::pts + j.  A node record holds a
// point's first cell U, the bits of its fractional offset fr in [0, 1) and of
// its modulated value (re, im) on one grid, so that pass 1 fetches a point
// with one 16-byte load.
struct PointArrays {
    int4* nd0;     // grid 0 (yw at f): node set 1, in bucket order
    int4* nd1;     // grid 1 (w at f): node set 1 (same U, fr as nd0)
    int4* nd2;     // grid 2 (w at 2f): node set 2
    float* w;      float* yw;      // prepared weights (direct method)
    double* dt;                    // t - tref (direct method)
};

struct PlanParams {
    int32_t m;
    int32_t mode;
    double reg_c;
    double reg_s;
    int32_t input_kind;
    int32_t method;
    int32_t kind;      // 0 = lomb, 1 = nfft
    int32_t window;    // GLS_WINDOW_*
    float es_beta;     // ES shape parameter
    int32_t fp32_coords;  // GLS_PRECISION_FP32: FP32 time offsets, cells and phases
    // ES: 1/psi_hat(xi) at xi = i / (2 kEsTab), i = 0 .. kEsTab + 1 (the
    // library's resident device copy for (m, beta), es_table_device)
    const float* es_deconv;
};

constexpr int kEsTab = 16384;   // ES deconvolution table intervals over xi in [0, 1/2]

// tiles of a size bucket (n1, p2): the caller's override (gls_tile_set), else
// the measured table compiled in, else the default rules (tiles.hpp)
struct BucketTiles {
    int c1, cr;
};
// whether this library has a kernel for the tile (tiles_gen.hpp lists them)
bool pass1_tile_built(int n1, int c1);
bool pass2_tile_built(int p2, int cr);

// launchers (defined in the kernel translation units)
// curves of more than kWidePrep points are prepared by 1024-thread blocks
constexpr int kWidePrep = 16384;
// For the NFFT method their point bucketing is spread over many blocks of
// kWideChunk points each (a counting pass and a scatter pass), with this
// per-curve record in the workspace: the raw-input weight constants found by
// the one-block reductions, the count of counting blocks done, and per node
// set the bucket counts, turned into bucket cursors by the last counting block.
constexpr int kWideChunk = 2048;
constexpr int kMaxBuckets = 4096;   // n2 / c1 <= 2^12
constexpr double kLowCycles = 3.0;
constexpr int kLowMax = 256;
constexpr int kLowSplit = 8;
// (the record also holds, per low frequency, the six FP64 sums of each of the
// kLowSplit workgroups of the low-frequency kernel, which the last of them adds
// in slice order, and how many of them are done, cleared at every execute)
struct WideAux {
    double ymean;
    double s0inv;
    int32_t done;       // counting blocks finished
    int32_t red_done;   // reduction blocks finished
    int32_t yy_done;    // scatter blocks finished (variance partials)
    int32_t pad;
    int32_t hist[2][kMaxBuckets];
    double low[kLowMax][kLowSplit][6];   // per low frequency and slice: the six sums
    int32_t low_done[kLowMax];
};
// Per-chunk partial reductions of a wide curve, one slot per chunk of
// kWideChunk points, combined in chunk order by the last block to finish (so
// the result does not depend on block timing).
struct WidePart {
    double tmin, tmax;   // time range
    double st, ct;       // sum and count of the finite t (FP32 coordinates only)
    double sy, cy;       // sum and count of the finite y (raw input)
    double sw, swy;      // sum of 1/dy^2 and of y/dy^2 (raw input)
    double syy;          // sum of (y - ymean - ybar)^2/dy^2 (raw input, from the scatter)
    double pad;
};
// aux: one WideAux per wide curve (NFFT method), or nullptr to prepare each
// wide curve in one block; part: ceil(wide_max_obs / kWideChunk) WidePart
// slots per wide curve; wide_max_obs: the largest n_obs in the wide list
hipError_t launch_prep(const CurveGeom* d_geom, const int32_t* small, int nsmall,
                       const int32_t* wide, int nwide, CurveState* state,
                       const double* t, const void* a, const void* b,
                       PointArrays pa, int32_t* bkt, const PlanParams& pp,
                       int y_is_f32, WideAux* aux, WidePart* part, int wide_max_obs,
                       hipStream_t s);
// work-items of one launch of the NFFT passes: the dispatch counts a grid's
// work-items in 32 bits, so a size bucket with more goes out in several
// launches over consecutive curves (max_launch_items(): this, or a smaller
// cap set by gls_launch_items_set)
constexpr int64_t kMaxLaunchItems = (int64_t)1 << 31;
int64_t max_launch_items();
// pass 1 stores a size bucket's grids non-temporally when they exceed this
// (the memory-side cache holds 256 MB)
constexpr int64_t kStreamStoreBytes = (int64_t)96 << 20;
// curves of one plan: the per-curve launches (preparation, low-frequency
// sums, top-k: at most 512 threads per curve) stay within kMaxLaunchItems
constexpr int kMaxPlanCurves = 1 << 22;
// GLS_SPREAD_AUTO spreads a curve in fixed point when it has at least one
// point per kFixedCellsPerPoint grid cells and its pass-1 columns hold at
// least kFixedMinColumn cells (on shorter columns the per-row counting and
// conversion cost more than the integer adds save)
constexpr int kFixedCellsPerPoint = 80;
constexpr int kFixedMinColumn = 64;
// pass 1 and 2 of one size bucket: bgeom holds the geometry of its ncurves
// curves, one after another (a block reads its curve's entry directly)
// (c1: the bucket's pass-1 tile, cr: its pass-2 tile; fix: fixed-point
// spreading; nts: non-temporal grid stores); pass 1's bkt points at the first
// curve's bucket offsets, which the others' follow, 2 (nb + 1) each
hipError_t launch_pass1(int n1, int c1, bool fix, bool nts, const CurveGeom* bgeom, int ncurves,
                        int ngrids, int p2, PointArrays pa, const int32_t* bkt, float2* grids,
                        const PlanParams& pp, hipStream_t s);
hipError_t launch_pass2_lomb(int p2, int cr, int n1, const CurveGeom* bgeom,
                             const CurveState* state, int ncurves, const float2* grids,
                             const PlanParams& pp, void* power, int out_f64,
                             hipStream_t s);
hipError_t launch_pass2_nfft(int p2, int cr, int n1, const CurveGeom* bgeom,
                             const CurveState* state, int ncurves, const float2* grids,
                             const PlanParams& pp, float2* out, hipStream_t s);
hipError_t launch_dirsum(const CurveGeom* d_geom, const CurveState* state,
                         const int32_t* tile_curve, const int32_t* tile_first,
                         int ntiles, int rf, PointArrays pa,
                         const PlanParams& pp, void* power, int out_f64,
                         hipStream_t s);
// direct FP64 sums for a curve's lowest frequencies, f_k (t_max - t_min) <
// kLowCycles (all of them), overwriting the NFFT powers there; with aux (one
// WideAux per wide curve) a wide curve's first kLowMax such frequencies are
// each summed by kLowSplit workgroups, any further ones by one workgroup
hipError_t launch_lowfreq(const CurveGeom* d_geom, const int32_t* small, int nsmall,
                          const int32_t* wide, int nwide, const CurveState* state,
                          PointArrays pa, const PlanParams& pp, void* power, int out_f64,
                          WideAux* aux, hipStream_t s);
hipError_t launch_dirsum_custom(const CurveGeom* d_geom, const CurveState* state,
                                const int32_t* tile_curve, const int32_t* tile_first,
                                int ntiles, PointArrays pa, const PlanParams& pp,
                                const double* freqs, void* power, int out_f64,
                                hipStream_t s);
hipError_t launch_gaussian_grid(const double* t, const double* y, int n_obs,
                                int ng, int m, double b, double tmin,
                                double tmax, double spp, float* grid,
                                hipStream_t s);
hipError_t launch_topk(const CurveGeom* d_geom, int ncurves, const float* power,
                       const uint8_t* mask, int k, float* top_power,
                       int32_t* top_index, hipStream_t s);

}  // namespace gls
