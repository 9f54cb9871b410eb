// Internal declarations shared by the libgls.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gls.h"

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

// ---------------------------------------------------------------------------
// resident twiddle tables (built by gls_init, one set per device)
//   hi[i] = exp(+2 pi i * i / 4096),      i < 4096
//   lo[i] = exp(+2 pi i * i / 2^24),      i < 4096
// exp(2 pi i q / 2^p), q < 2^p <= 2^24:  Q = q << (24-p),
//   = hi[Q >> 12] * lo[Q & 4095]
// ---------------------------------------------------------------------------
struct Twiddles {
    const float2* hi;
    const float2* lo;
};
Twiddles device_twiddles();

constexpr int kMaxLog2Grid = 24;   // largest grid n = 2^24 cells
constexpr int kMinLog2Grid = 10;   // smallest grid n = 2^10 cells

// ---------------------------------------------------------------------------
// per-curve plan data (host computed, uploaded once per plan)
// ---------------------------------------------------------------------------
struct CurveGeom {
    int64_t obs;       // first observation index in t / a / b
    int64_t out;       // first output element
    int64_t pts;       // first point slot in the per-point workspace arrays
    int64_t bkt;       // first int of this curve's bucket table (2*(nb+1))
    int64_t grid;      // first float2 of this curve's 3 grids (3*n)
    double f0;
    double df;
    double b;          // Gaussian width in cells^2
    double ybar_in;    // GLS_INPUT_WEIGHTED
    double yy_in;      // GLS_INPUT_WEIGHTED
    int32_t n_obs;
    int32_t nf;        // K, number of output frequencies
    int32_t p;         // log2 of grid size n
    int32_t p1;        // log2 of pass-1 column length n1
    int32_t c1;        // pass-1 columns per tile
    int32_t kc;        // K/2, band centre index
};

// written by the prep kernel
struct CurveState {
    double tref;       // time origin (min t)
    double ybar;
    double yy;
    double tspan;      // max t - min t
};

// per-point workspace (structure of arrays, indexed by CurveGeom::pts + j)
struct PointArrays {
    int32_t* U1;   float* fr1;     // grid type 1 (SYW, W1): sorted by bucket
    float2* v0;    float2* v1;     // modulated yw, w
    int32_t* U2;   float* fr2;     // grid type 2 (W2 at 2f)
    float2* v2;                    // modulated w at 2f
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
};

// launchers (defined in the kernel translation units)
hipError_t launch_prep(const CurveGeom* d_geom, int ncurves, CurveState* state,
                       const double* t, const void* a, const void* b,
                       PointArrays pa, int32_t* bkt, const PlanParams& pp,
                       int y_is_f32, hipStream_t s);
hipError_t launch_pass1(int p, int p1, int c1, const CurveGeom* d_geom,
                        const int32_t* curves, int ncurves, int ngrids,
                        PointArrays pa, const int32_t* bkt, float2* grids,
                        int m, Twiddles tw, hipStream_t s);
hipError_t launch_pass2_lomb(int p, int p1, const CurveGeom* d_geom,
                             const CurveState* state, const int32_t* curves,
                             int ncurves, const float2* grids,
                             const PlanParams& pp, void* power, int out_f64,
                             Twiddles tw, hipStream_t s);
hipError_t launch_pass2_nfft(int p, int p1, const CurveGeom* d_geom,
                             const CurveState* state, const int32_t* curves,
                             int ncurves, const float2* grids, float2* out,
                             Twiddles tw, hipStream_t s);
hipError_t launch_dirsum(const CurveGeom* d_geom, const CurveState* state,
                         const int32_t* tile_curve, const int32_t* tile_first,
                         int ntiles, int rf, PointArrays pa,
                         const PlanParams& pp, void* power, int out_f64,
                         hipStream_t s);
hipError_t launch_gaussian_grid(const double* t, const double* y, int n_obs,
                                int ng, int m, double b, double tmin,
                                double tmax, double spp, float* grid,
                                hipStream_t s);
hipError_t launch_topk(const CurveGeom* d_geom, int ncurves, const float* power,
                       const uint8_t* mask, int k, float* top_power,
                       int32_t* top_index, hipStream_t s);

// size rules shared by planner and kernels
inline int pass1_log2_len(int p) {
    // n = n1 * n2, n1 = 2^p1 column length (pass 1), n2 = 2^(p-p1) row length
    int p1 = p - 8;
    if (p1 > 12) p1 = 12;
    if (p1 < 4) p1 = 4;
    if (p - p1 < 6) p1 = p - 6;
    return p1;
}
constexpr int pass1_cols(int p1) {
    int c = 8192 >> p1;
    if (c > 16) c = 16;
    if (c < 1) c = 1;
    return c;
}
constexpr int pass2_rows(int p2) {
    int c = 2048 >> p2;
    if (c > 16) c = 16;
    if (c < 1) c = 1;
    return c;
}

}  // namespace gls
