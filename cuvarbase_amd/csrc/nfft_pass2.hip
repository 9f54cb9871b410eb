// Pass 2 of the batched NFFT-GLS pipeline: row FFTs of the four-step split,
// the window deconvolution and the GLS power assembly (or the NFFT output).
//
// Grid of n = n1*n2 cells (n2 = 2^p2); pass 1 has left A[k1,u2] in row blocks
// of R = kRowBlock rows (A[(k1/R) R n2 + u2 R + k1 % R]) with the inter-pass
// twiddle applied, so row k1 of the transform is
// G[k1 + n1 k2] = sum_u2 A[k1,u2] exp(2 pi i u2 k2/n2).  A
// workgroup transforms CR rows of the three grids of a curve (SYW, SW(f),
// SW(2f) -- all indexed by the same output k), applies the deconvolution
// (normalize, cuvarbase/kernels/cunfft.cu:243-245) and forms the GLS power
// (lspow, cuvarbase/kernels/lomb.cu:126-141, in closed form) without ever
// writing the transforms.
#include <algorithm>
#include <type_traits>

#include "device_util.hpp"
#include "lds_fft.hpp"

namespace gls {

namespace {

template <int V>
using Ic = std::integral_constant<int, V>;

// OUT: 0 = f32 power, 1 = f64 power, 2 = complex64 NFFT (single grid);
// CR rows of each grid per workgroup
template <int P2, int CR, int OUT>
__global__ __launch_bounds__(pass2_threads_for(P2, CR, OUT)) void pass2_kernel(
    const CurveGeom* __restrict__ bgeom, const CurveState* __restrict__ state, int n1,
    const float2* __restrict__ grids, PlanParams pp, void* __restrict__ out) {
    constexpr int NT = pass2_threads_for(P2, CR, OUT);
    constexpr int N2 = 1 << P2;
    constexpr int NG = (OUT == 2) ? 1 : 3;
    constexpr int B = NG * CR;
    constexpr int BS = lds_stride(B);
    __shared__ float2 tile[lds_size(N2, BS)];
    __shared__ TwTable<N2> T;
    const int tid = threadIdx.x;
    const int nrb = (n1 + CR - 1) / CR;   // the last row block may be partial
    const int bid = xcd_logical_block(blockIdx.x, gridDim.x);
    const int rb = bid % nrb;
    const CurveGeom& G = bgeom[bid / nrb];
    const int ci = G.ci;
    const int n = G.n;
    const int r0 = rb * CR;

    // row FFTs of the tile (rows r0 .. r0+CR-1 of the three grids); the first
    // stage loads its inputs straight from HBM
    const uint32_t gs = (uint32_t)grid_cells(n1, P2);   // cells of one grid
    const __amdgpu_buffer_rsrc_t src = grid_rsrc(grids + G.grid, 8u * (NG * gs));
    build_twiddles<N2, NT>(T, tid);
    lds_fft_from<N2, B, BS, NT>(tile, T, tid, [&](int j, int q, int lr, int bb) -> float2 {
        // column u2 = j + q lr of row `row` in the row-block layout.  Rows past
        // n1 (a partial last block) read row n1 - 1 instead; their transforms
        // are never emitted
        constexpr int R = kRowBlock;
        const int gi = bb / CR;
        const int row = min(r0 + bb - gi * CR, n1 - 1);
        return grid_load(src, 8u * (gi * gs + (uint32_t)((row / R) * (R * N2) + row % R)) +
                                  8u * (uint32_t)((j + q * lr) * R));
    });

    const int K = G.nf;
    const int kc = G.kc;
    const float bpi2 = G.g_bpi2;
    const float ninv = G.ninv;
    const CurveState st = state[ci];
    const float yy32 = (float)st.yy, ybar32 = (float)st.ybar;
    const float rc32 = (float)pp.reg_c, rs32 = (float)pp.reg_s;
    // this curve's FP32 powers (OUT == 0)
    float* const dst32 = (OUT == 0) ? (float*)out + G.out : nullptr;
    // output slots i = tid + q NT: the row r = i % CR is the thread's own
    // (NT is a multiple of CR), k2 = i / CR steps by NT / CR, and the three
    // grids' LDS places step by a fixed amount
    static_assert(NT % CR == 0, "a thread's output slots share one row");
    constexpr int KSTEP = NT / CR;
    constexpr int QMAX = (N2 + KSTEP - 1) / KSTEP;
    const int r = tid % CR;
    const int k20 = tid / CR;
    const LdsRun<BS, KSTEP> p0(k20, r), p1(k20, CR + r), p2(k20, 2 * CR + r);
    if (r0 + r >= n1) return;   // a row past n1 (no barrier follows)
    const int l0 = r0 + r + n1 * k20;
    const int lstep = n1 * KSTEP;
    // byte-addressed views: the three runs' LDS places (linear in q, so the
    // slot offsets are immediates), the ES table (kEsTab + 2 entries) and this
    // curve's K power outputs
    static_assert(LdsRun<BS, KSTEP>::LINEAR, "pass-2 output runs are linear");
    constexpr int DQ = LdsRun<BS, KSTEP>::DELTA;
    const float2* const t0p = tile + p0.at(0);
    const float2* const t1p = tile + p1.at(0);
    const float2* const t2p = tile + p2.at(0);
    const __amdgpu_buffer_rsrc_t es_tab = grid_rsrc(pp.es_deconv, 4u * (kEsTab + 2));
    const __amdgpu_buffer_rsrc_t pw = grid_rsrc(dst32, (OUT == 0) ? 4u * (uint32_t)K : 0u);
    const float den_std = yy32 + ybar32 * ybar32;
    // the slot loop, specialised on the power mode and the window (both
    // uniform per launch) and unrolled so that the LDS places are immediates
    auto outputs = [&](auto mode_c, auto es_c) {
        constexpr int MODE = decltype(mode_c)::value;
        constexpr bool ES = decltype(es_c)::value;
#pragma unroll 8
        for (int q = 0; q < QMAX; ++q) {
            if (k20 + q * KSTEP >= N2) break;
            const int l = l0 + q * lstep;
            int lp;
            if (l < K - kc) lp = l;
            else if (l >= n - kc) lp = l - n;
            else continue;
            const float xi = (float)lp * ninv;
            float dc;
            if constexpr (ES) {
                // 1/psi_hat(|xi|) by linear interpolation in the plan's table
                const float u = fabsf(xi) * (2.0f * kEsTab);
                const int i0 = min((int)u, kEsTab);
                const float f = u - (float)i0;
                const float2 tt = grid_load(es_tab, 4u * (uint32_t)i0);
                dc = tt.x + f * (tt.y - tt.x);
            } else {
                dc = expf(bpi2 * xi * xi);
            }
            const uint32_t k = (uint32_t)(lp + kc);
            if constexpr (OUT == 2) {
                const float2 V = cscale(t0p[q * DQ], dc);
                // restore the absolute time origin: exp(2 pi i f_k tref)
                const double fk = G.f0 + (double)k * G.df;
                const float2 ph = expi_rev(fk * st.tref);
                ((float2*)out + G.out)[k] = cmul(V, ph);
            } else {
                const float2 syw = t0p[q * DQ];
                const float2 w1 = t1p[q * DQ];
                const float2 w2 = t2p[q * DQ];
                if constexpr (OUT == 0) {
                    buf_store_f32(pw, 4u * k,
                                  gls_power_f32(dc * w1.x, dc * w1.y, dc * w2.x, dc * w2.y,
                                                dc * syw.x, dc * syw.y, yy32, ybar32, MODE,
                                                rc32, rs32, den_std));
                } else {
                    // FP64 power output: the assembly in FP64 as well
                    const double d = (double)dc;
                    ((double*)out + G.out)[k] =
                        gls_power(d * w1.x, d * w1.y, d * w2.x, d * w2.y, d * syw.x, d * syw.y,
                                  st.yy, st.ybar, pp.mode, pp.reg_c, pp.reg_s);
                }
            }
        }
    };
    using ES_ = std::true_type;
    using GA_ = std::false_type;
    if constexpr (OUT == 0) {
        const int mode = pp.mode;
        if (pp.window == GLS_WINDOW_ES) {
            if (mode == GLS_MODE_FLOATING_MEAN) outputs(Ic<GLS_MODE_FLOATING_MEAN>{}, ES_{});
            else if (mode == GLS_MODE_STANDARD) outputs(Ic<GLS_MODE_STANDARD>{}, ES_{});
            else if (mode == GLS_MODE_WINDOW) outputs(Ic<GLS_MODE_WINDOW>{}, ES_{});
            else outputs(Ic<-1>{}, ES_{});
        } else {
            if (mode == GLS_MODE_FLOATING_MEAN) outputs(Ic<GLS_MODE_FLOATING_MEAN>{}, GA_{});
            else if (mode == GLS_MODE_STANDARD) outputs(Ic<GLS_MODE_STANDARD>{}, GA_{});
            else if (mode == GLS_MODE_WINDOW) outputs(Ic<GLS_MODE_WINDOW>{}, GA_{});
            else outputs(Ic<-1>{}, GA_{});
        }
    } else {
        if (pp.window == GLS_WINDOW_ES) outputs(Ic<-1>{}, ES_{});
        else outputs(Ic<-1>{}, GA_{});
    }
}

template <int P2, int CR, int OUT>
hipError_t pass2_dispatch(int n1, const CurveGeom* bgeom, const CurveState* state, int ncurves,
                          const float2* grids, const PlanParams& pp, void* out, hipStream_t s) {
    constexpr int NT = pass2_threads_for(P2, CR, OUT);
    const int64_t per_curve = (int64_t)((n1 + CR - 1) / CR) * NT;
    // curves per launch: blocks x threads within max_launch_items()
    const int chunk = (int)std::min<int64_t>(ncurves, std::max<int64_t>(1, max_launch_items() / per_curve));
    for (int c0 = 0; c0 < ncurves; c0 += chunk) {
        const int cnt = std::min(chunk, ncurves - c0);
        hipLaunchKernelGGL((pass2_kernel<P2, CR, OUT>),
                           dim3((unsigned)((int64_t)cnt * ((n1 + CR - 1) / CR))), dim3(NT), 0, s,
                           bgeom + c0, state, n1, grids, pp, out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// the (p2, cr) pairs this library has kernels for
template <int OUT>
hipError_t pass2_switch(int p2, int cr, int n1, const CurveGeom* bgeom, const CurveState* state,
                        int ncurves, const float2* grids, const PlanParams& pp, void* out,
                        hipStream_t s) {
#define GLS_P2_CASE(P, C)                                                              \
    if (p2 == P && cr == C)                                                            \
        return pass2_dispatch<P, C, OUT>(n1, bgeom, state, ncurves, grids, pp, out, s);
    GLS_PASS2_TILES(GLS_P2_CASE)
#undef GLS_P2_CASE
    return hipErrorInvalidValue;
}

}  // namespace

bool pass2_tile_built(int p2, int cr) {
#define GLS_P2_HAS(P, C) \
    if (p2 == P && cr == C) return true;
    GLS_PASS2_TILES(GLS_P2_HAS)
#undef GLS_P2_HAS
    return false;
}

hipError_t launch_pass2_lomb(int p2, int cr, int n1, const CurveGeom* bgeom,
                             const CurveState* state, int ncurves, const float2* grids,
                             const PlanParams& pp, void* power, int out_f64, hipStream_t s) {
    if (ncurves <= 0) return hipSuccess;
    if (out_f64) return pass2_switch<1>(p2, cr, n1, bgeom, state, ncurves, grids, pp, power, s);
    return pass2_switch<0>(p2, cr, n1, bgeom, state, ncurves, grids, pp, power, s);
}

hipError_t launch_pass2_nfft(int p2, int cr, int n1, const CurveGeom* bgeom,
                             const CurveState* state, int ncurves, const float2* grids,
                             const PlanParams& pp, float2* out, hipStream_t s) {
    if (ncurves <= 0) return hipSuccess;
    return pass2_switch<2>(p2, cr, n1, bgeom, state, ncurves, grids, pp, out, s);
}

}  // namespace gls
