// Pass 1 of the batched NFFT-GLS pipeline: spreading, column FFTs and the
// inter-pass twiddle (pass 2, the row FFTs and the power assembly, is in
// nfft_pass2.hip).
//
// Grid of n = n1*n2 cells (n2 = 2^p2), cell u = n2*u1 + u2 (u1 row, u2
// column).  The unnormalised + sign transform G[l] = sum_u g[u] exp(2 pi i l u/n)
// is done four-step:
//   pass 1  per column u2:  A[k1,u2] = exp(2 pi i u2 k1/n) sum_u1 g[n2 u1+u2] exp(2 pi i u1 k1/n1)
//   pass 2  per row k1:     G[k1 + n1 k2] = sum_u2 A[k1,u2] exp(2 pi i u2 k2/n2)
// Pass 1 builds its columns directly in LDS by window spreading (the work of
// fast_gaussian_grid, cuvarbase/kernels/cunfft.cu:115-163) from the points the
// prep kernel bucketed by column block, so the spread grid never touches HBM.
// A is stored in row blocks of R = kRowBlock rows, A[(k1/R) R n2 + u2 R + k1 % R]
// (gls_internal.hpp), so a column goes out as segments of R cells and a
// pass-2 tile of R rows is one contiguous run.
#include <algorithm>
#include <type_traits>

#include "device_util.hpp"
#include "lds_fft.hpp"

namespace gls {

namespace {

// threads of a pass-1 workgroup: 512 for tiles of 6144 cells or more (one
// radix-16 butterfly of the first stage per thread at 8192 cells; two
// columns of the long lengths), 256 below
constexpr int pass1_threads(int n1, int c1) { return n1 * c1 >= 6144 ? 512 : 256; }

// ---------------------------------------------------------------------------
// pass 1: spread + column FFT + inter-pass twiddle
// ---------------------------------------------------------------------------
// FIX: fixed-point spreading (GLS_SPREAD_FIXED): the tile holds int32 pairs
// while the taps are added, each tap rounded to a multiple of 2^-e and added
// as one 64-bit integer im 2^32 + re, then is converted to float2 in place
// before the FFT.  All the cells a point touches in a tile lie in one row (its
// columns are consecutive and do not wrap), so e is chosen per row: from the
// number of the tile's points in the row times their largest |component|
// (window weights are at most 1), no cell of the row can exceed 2^30.  Counts
// and maxima are integer LDS atomics, so e, the roundings and the sums do not
// depend on the order of the points.
// C1: grid columns per workgroup (one LDS tile of N1 x C1 cells)
// NTS: the grids are stored non-temporally (launches whose grids are larger
// than the memory-side cache; pass 2 reads them from HBM either way)
template <int N1, int C1, bool FIX, bool NTS>
__global__ __launch_bounds__(pass1_threads(N1, C1)) void pass1_kernel(const CurveGeom* __restrict__ bgeom,
                                                    int ngrids, int p2, PointArrays pa,
                                                    const int32_t* __restrict__ bkt,
                                                    float2* __restrict__ grids, int m,
                                                    int window, float es_beta, float ainv) {
    static_assert(pass1_tile_ok(N1, C1), "pass-1 tile out of range");
    constexpr int NT1 = pass1_threads(N1, C1);
    constexpr int BS = lds_stride(C1);
    constexpr int NH = (N1 + 63) / 64;
    // (rounded up to whole 16-byte pairs so that it is cleared with 128-bit stores)
    constexpr int TL = (lds_size(N1, BS) + 1) & ~1;
    __shared__ __align__(16) float2 tile[TL];
    __shared__ TwTable<N1> T;
    __shared__ float2 ehi[C1 * NH];   // exp(2 pi i u2 64 kh / n)
    __shared__ float2 elo[C1 * 64];   // exp(2 pi i u2 kl / n)
    __shared__ signed char rexp[FIX ? N1 : 1];   // fixed-point exponent per row
    const int tid = threadIdx.x;
    const int n2 = 1 << p2;
    const int nb = n2 / C1;
    const int bid = xcd_logical_block(blockIdx.x, gridDim.x);
    const int jb = bid % nb;
    const int gi = (bid / nb) % ngrids;
    const int ce = bid / (nb * ngrids);   // the curve's place in the bucket entry
    const CurveGeom& G = bgeom[ce];
    const int n = G.n;
    const int a = jb * C1;

    // point set: buckets jb-nbk .. jb (cyclic), at most two contiguous ranges.
    // The entry's curves keep their bucket offsets in order, 2 (nb + 1) each,
    // so these loads need not wait for the curve's record
    const int g = (gi == 2) ? 2 : 1;
    const int32_t* bk = bkt + (int64_t)ce * (2 * (nb + 1)) + (int64_t)(g - 1) * (nb + 1);
    const int nbk = (2 * m + C1 - 1) / C1;
    int s0, e0, s1, e1;
    const int first = jb - nbk;
    if (first >= 0) {
        s0 = bk[first];
        e0 = bk[jb + 1];
        s1 = e1 = 0;
    } else {
        s0 = bk[nb + first];
        e0 = bk[nb];
        s1 = bk[0];
        e1 = bk[jb + 1];
    }
    const int len0 = e0 - s0;
    const int total = len0 + (e1 - s1);
    // node records {U, fr, value} of this grid, one 16-byte load per point
    const int4* Na = (gi == 0) ? pa.nd0 : (gi == 1 ? pa.nd1 : pa.nd2);

    // the first PF points of each thread are loaded before the tile is
    // cleared and the twiddle tables are built, so their latency overlaps
    // that work
    constexpr int PF = 2;
    int4 pn[PF];
#pragma unroll
    for (int f = 0; f < PF; ++f) {
        const int q = tid + f * NT1;
        pn[f] = make_int4(0, 0, 0, 0);
        if (q < total) pn[f] = Na[G.pts + (q < len0 ? s0 + q : s1 + (q - len0))];
    }

    for (int i = tid; i < TL / 2; i += NT1)
        reinterpret_cast<float4*>(tile)[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    build_twiddles<N1, NT1>(T, tid);
    for (int i = tid; i < C1 * (NH + 64); i += NT1) {
        const int jj = i / (NH + 64);
        const int r = i - jj * (NH + 64);
        const int u2 = a + jj;
        // u2 < 2^12, r < 64, r - 64 < 2^6: the products stay below 2^24
        if (r < 64) elo[jj * 64 + r] = expi_ratio32((uint32_t)(u2 * r), n);
        else ehi[jj * NH + (r - 64)] = expi_ratio32((uint32_t)(u2 * 64 * (r - 64)), n);
    }
    __syncthreads();

    // window weight of a point at distance d (cells) from a cell:
    // Gaussian exp(-d^2/b)/sqrt(pi b), or ES exp(beta (sqrt(1 - z^2) - 1)),
    // z = d/(m + 1/2), zero for |z| >= 1
    const bool es = window == GLS_WINDOW_ES;
    const float binv = G.g_binv;   // log2(e)/b
    const float norm = G.g_norm;
    const float bl2 = es_beta * 1.4426950408889634f;          // beta log2(e)
    const int two_m = 2 * m;
    // the tile row of a point's cells (the first in-tile tap is at column a,
    // or, when the point's taps start inside the tile, at its first cell U)
    auto row_of = [&](int U) {
        const int k0 = (a - (U & (n2 - 1))) & (n2 - 1);
        int cell = U + (k0 <= two_m ? k0 : 0);
        cell = cell < 0 ? cell + n : (cell >= n ? cell - n : cell);
        return cell >> p2;
    };
    if (FIX) {
        // per row: point count (first N1 ints of the cleared tile) and the
        // largest |component| as non-negative float bits (next N1 ints)
        int* cnt = reinterpret_cast<int*>(tile);
        int* mxb = cnt + N1;
        auto count = [&](int4 nd) {
            const int r = row_of(nd.x);
            atomicAdd(&cnt[r], 1);
            atomicMax(&mxb[r], __float_as_int(fmaxf(fabsf(__int_as_float(nd.z)),
                                                    fabsf(__int_as_float(nd.w)))));
        };
#pragma unroll
        for (int f = 0; f < PF; ++f)
            if (tid + f * NT1 < total) count(pn[f]);
        for (int q = tid + PF * NT1; q < total; q += NT1)
            count(Na[G.pts + (q < len0 ? s0 + q : s1 + (q - len0))]);
        __syncthreads();
        for (int r = tid; r < N1; r += NT1)
            rexp[r] = (signed char)spread_exponent((float)cnt[r] * __int_as_float(mxb[r]));
        __syncthreads();
        for (int i = tid; i < N1; i += NT1)   // the 2 N1 ints = N1 float2 used above
            tile[i] = make_float2(0.f, 0.f);
        __syncthreads();
    }
    auto spread = [&](int4 nd) {
        const int U = nd.x;
        const float fr = __int_as_float(nd.y);
        float2 v = make_float2(__int_as_float(nd.z), __int_as_float(nd.w));
        if (FIX) {
            const float sc = ldexpf(1.0f, rexp[row_of(U)]);
            v = make_float2(v.x * sc, v.y * sc);
        }
        const int col0 = U & (n2 - 1);
#pragma unroll
        for (int jj = 0; jj < C1; ++jj) {
            const int k = (a + jj - col0) & (n2 - 1);
            if (k <= two_m) {
                int cell = U + k;
                cell = cell < 0 ? cell + n : (cell >= n ? cell - n : cell);
                const int u1 = cell >> p2;
                const float d = (float)(m - k) + fr;
                float wgt;
                if (es) {
                    const float z = d * ainv;
                    const float q = 1.0f - z * z;
                    wgt = q > 0.f ? exp2f(bl2 * (sqrtf(q) - 1.0f)) : 0.f;
                } else {
                    wgt = exp2f(-d * d * binv) * norm;
                }
                float2* cellp = &tile[lds_at<BS>(u1, jj)];
                if (FIX) {
                    // both components in one 64-bit integer add: the cell
                    // holds im 2^32 + re (two's complement), exact while
                    // |re|, |im| < 2^31, which the row exponent ensures
                    const long long re = __float2int_rn(v.x * wgt);
                    const long long im = __float2int_rn(v.y * wgt);
                    atomicAdd(reinterpret_cast<unsigned long long*>(cellp),
                              (unsigned long long)(im * 4294967296LL + re));
                } else {
                    atomicAdd(&cellp->x, v.x * wgt);
                    atomicAdd(&cellp->y, v.y * wgt);
                }
            }
        }
    };
#pragma unroll
    for (int f = 0; f < PF; ++f)
        if (tid + f * NT1 < total) spread(pn[f]);
    for (int q = tid + PF * NT1; q < total; q += NT1)
        spread(Na[G.pts + (q < len0 ? s0 + q : s1 + (q - len0))]);
    __syncthreads();
    if (FIX) {
        // integer cells -> float2 with their row's scale (exact powers of
        // two).  A cell's 64-bit sum im 2^32 + re keeps re in its low word
        // and im - 1 in its high word when re < 0
        constexpr int LC = ilog2_c(C1);
        for (int i = tid; i < N1 * C1; i += NT1) {
            float2* cp = &tile[i + (i >> 4)];   // lds_at<BS>(i / C1, i % C1)
            const int2 c = *reinterpret_cast<const int2*>(cp);
            const int im = c.y + (c.x < 0 ? 1 : 0);
            const float inv = ldexpf(1.0f, -(int)rexp[i >> LC]);
            *cp = make_float2((float)c.x * inv, (float)im * inv);
        }
        __syncthreads();
    }

    // column FFT; its last stage writes straight from registers to HBM with
    // the inter-pass twiddle exp(2 pi i u2 k1 / n) (consecutive lanes hold
    // consecutive k1 of a column, so the stores coalesce)
    const int64_t gs = grid_cells(N1, p2);   // cells of one grid
    const __amdgpu_buffer_rsrc_t out = grid_rsrc(grids + G.grid + gi * gs, 8u * (uint32_t)gs);
    const int rb = kRowBlock << p2;          // cells of one row block
    // (elo[0] = 1 exactly, so the product needs no special case; when the
    // output stride ns is a multiple of 64 a butterfly's outputs share one elo
    // entry and step through ehi, and the stride goes to the scalar offset)
    lds_fft_to<N1, C1, BS, NT1>(tile, T, tid, [&](int kb, int q, int ns, int jj, float2 v) {
        float2 w;
        if ((ns & 63) == 0) {
            w = cmul(ehi[jj * NH + (kb >> 6) + q * (ns >> 6)], elo[jj * 64 + (kb & 63)]);
        } else {
            const int k1 = kb + q * ns;
            w = cmul(ehi[jj * NH + (k1 >> 6)], elo[jj * 64 + (k1 & 63)]);
        }
        // row-block layout: when the output stride ns is a multiple of the
        // block height, a butterfly's outputs share their in-block offset and
        // step whole row blocks (the step goes to the scalar offset)
        constexpr int R = kRowBlock;
        constexpr int CP = NTS ? kStoreNonTemporal : 0;
        if ((ns % R) == 0) {
            grid_store<CP>(out, 8u * (uint32_t)((kb / R) * rb + (a + jj) * R + kb % R),
                           8u * (uint32_t)(q * (ns / R) * rb), cmul(v, w));
        } else {
            const int k1 = kb + q * ns;
            grid_store<CP>(out, 8u * (uint32_t)((k1 / R) * rb + (a + jj) * R + k1 % R), 0u,
                           cmul(v, w));
        }
    });
}

template <int N1, int C1>
hipError_t pass1_dispatch(bool fix, bool nts, const CurveGeom* bgeom, int ncurves, int ngrids,
                          int p2,
                          PointArrays pa, const int32_t* bkt, float2* grids,
                          const PlanParams& pp, hipStream_t s) {
    constexpr int NT = pass1_threads(N1, C1);
    const int nb = (1 << p2) / C1;
    if (nb < 1 || ngrids < 1) return hipErrorInvalidValue;
    const float ainv = 1.0f / ((float)pp.m + 0.5f);
    // curves per launch: blocks x threads within max_launch_items()
    const int64_t per_curve = (int64_t)ngrids * nb * NT;
    const int chunk = (int)std::min<int64_t>(ncurves, std::max<int64_t>(1, max_launch_items() / per_curve));
    for (int c0 = 0; c0 < ncurves; c0 += chunk) {
        const int cnt = std::min(chunk, ncurves - c0);
        const dim3 grid((unsigned)((int64_t)cnt * ngrids * nb));
        // a curve's bucket offsets follow its predecessors', 2 (nb + 1) each
        const int32_t* bk = bkt + (int64_t)c0 * 2 * (nb + 1);
#define GLS_P1_LAUNCH(F, S)                                                                    \
    hipLaunchKernelGGL((pass1_kernel<N1, C1, F, S>), grid, dim3(NT), 0, s, bgeom + c0, ngrids, \
                       p2, pa, bk, grids, pp.m, pp.window, pp.es_beta, ainv)
        if (fix) {
            if (nts) GLS_P1_LAUNCH(true, true);
            else GLS_P1_LAUNCH(true, false);
        } else {
            if (nts) GLS_P1_LAUNCH(false, true);
            else GLS_P1_LAUNCH(false, false);
        }
#undef GLS_P1_LAUNCH
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace

bool pass1_tile_built(int n1, int c1) {
#define GLS_P1_HAS(N, C) \
    if (n1 == N && c1 == C) return true;
    GLS_PASS1_TILES(GLS_P1_HAS)
#undef GLS_P1_HAS
    return false;
}

// the (n1, c1) pairs this library has kernels for
hipError_t launch_pass1(int n1, int c1, bool fix, bool nts, const CurveGeom* bgeom, int ncurves,
                        int ngrids, int p2, PointArrays pa, const int32_t* bkt, float2* grids,
                        const PlanParams& pp, hipStream_t s) {
    if (ncurves <= 0) return hipSuccess;
#define GLS_P1_CASE(N, C)  \
    if (n1 == N && c1 == C) \
        return pass1_dispatch<N, C>(fix, nts, bgeom, ncurves, ngrids, p2, pa, bkt, grids, pp, s);
    GLS_PASS1_TILES(GLS_P1_CASE)
#undef GLS_P1_CASE
    return hipErrorInvalidValue;
}

}  // namespace gls
