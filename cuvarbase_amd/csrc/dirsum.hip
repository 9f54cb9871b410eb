// Direct-summation GLS (use_fft=False).
//
// Replaces lomb_dirsum (cuvarbase/kernels/lomb.cu:144-165): six O(N) sums per
// frequency, then the power.  Each thread owns RF consecutive frequencies of
// one curve; observations are staged through LDS in chunks of 256.  Per
// (observation, thread) the phase of the first frequency is reduced mod 1 in
// FP64 and evaluated once in FP32; the next RF-1 frequencies follow by
// rotation with exp(2 pi i df (t - tref)) (one complex multiply each).  Sums
// are accumulated in FP32 within a chunk and in FP64 across chunks.
#include "device_util.hpp"
#include "lds_fft.hpp"

namespace gls {

namespace {

constexpr int NT = 256;

template <int RF, int OUT>
__global__ __launch_bounds__(NT) void dirsum_kernel(const CurveGeom* __restrict__ geom,
                                                    const CurveState* __restrict__ state,
                                                    const int32_t* __restrict__ tile_curve,
                                                    const int32_t* __restrict__ tile_first,
                                                    PointArrays pa, PlanParams pp,
                                                    void* __restrict__ out) {
    __shared__ double sdt[NT];
    __shared__ float sw[NT];
    __shared__ float syw[NT];
    __shared__ float2 srot[NT];
    const int tid = threadIdx.x;
    const int ci = tile_curve[blockIdx.x];
    const CurveGeom& G = geom[ci];
    const int N = G.n_obs;
    const int k0 = tile_first[blockIdx.x] + tid * RF;
    const double fk0 = G.f0 + (double)k0 * G.df;

    double acc[6][RF];
#pragma unroll
    for (int q = 0; q < 6; ++q)
#pragma unroll
        for (int r = 0; r < RF; ++r) acc[q][r] = 0.0;

    for (int base = 0; base < N; base += NT) {
        const int j = base + tid;
        if (j < N) {
            const double dt = pa.dt[G.pts + j];
            sdt[tid] = dt;
            sw[tid] = pa.w[G.pts + j];
            syw[tid] = pa.yw[G.pts + j];
            srot[tid] = expi_rev(G.df * dt);
        }
        __syncthreads();
        const int cnt = min(NT, N - base);
        float a[6][RF];
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int r = 0; r < RF; ++r) a[q][r] = 0.f;
        for (int jj = 0; jj < cnt; ++jj) {
            const double dt = sdt[jj];
            const float w = sw[jj];
            const float yw = syw[jj];
            const float2 rot = srot[jj];
            float2 z = expi_rev(fk0 * dt);
#pragma unroll
            for (int r = 0; r < RF; ++r) {
                const float c = z.x, s = z.y;
                a[0][r] += w * c;
                a[1][r] += w * s;
                a[2][r] += w * (c * c - s * s);
                a[3][r] += w * (2.f * c * s);
                a[4][r] += yw * c;
                a[5][r] += yw * s;
                if (r + 1 < RF) z = cmul(z, rot);
            }
        }
#pragma unroll
        for (int q = 0; q < 6; ++q)
#pragma unroll
            for (int r = 0; r < RF; ++r) acc[q][r] += (double)a[q][r];
        __syncthreads();
    }

    const CurveState st = state[ci];
#pragma unroll
    for (int r = 0; r < RF; ++r) {
        const int k = k0 + r;
        if (k < G.nf) {
            const double P = gls_power(acc[0][r], acc[1][r], acc[2][r], acc[3][r], acc[4][r],
                                          acc[5][r], st.yy, st.ybar, pp.mode, pp.reg_c,
                                          pp.reg_s);
            if (OUT == 0) ((float*)out)[G.out + k] = (float)P;
            else ((double*)out)[G.out + k] = P;
        }
    }
}

// Direct sums at explicit frequencies freqs[out + k] (any spacing): the work of
// lomb_dirsum_custom_frq (cuvarbase/kernels/lomb.cu:167-187).  One frequency
// per thread, each phase reduced mod 1 in FP64.
template <int OUT>
__global__ __launch_bounds__(NT) void dirsum_custom_kernel(const CurveGeom* __restrict__ geom,
                                                           const CurveState* __restrict__ state,
                                                           const int32_t* __restrict__ tile_curve,
                                                           const int32_t* __restrict__ tile_first,
                                                           PointArrays pa, PlanParams pp,
                                                           const double* __restrict__ freqs,
                                                           void* __restrict__ out) {
    __shared__ double sdt[NT];
    __shared__ float sw[NT];
    __shared__ float syw[NT];
    const int tid = threadIdx.x;
    const int ci = tile_curve[blockIdx.x];
    const CurveGeom& G = geom[ci];
    const int N = G.n_obs;
    const int k = tile_first[blockIdx.x] + tid;
    const double f = (k < G.nf) ? freqs[G.out + k] : 0.0;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int base = 0; base < N; base += NT) {
        const int j = base + tid;
        if (j < N) {
            sdt[tid] = pa.dt[G.pts + j];
            sw[tid] = pa.w[G.pts + j];
            syw[tid] = pa.yw[G.pts + j];
        }
        __syncthreads();
        const int cnt = min(NT, N - base);
        float a[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int jj = 0; jj < cnt; ++jj) {
            const float2 z = expi_rev(f * sdt[jj]);
            const float w = sw[jj], yw = syw[jj];
            a[0] += w * z.x;
            a[1] += w * z.y;
            a[2] += w * (z.x * z.x - z.y * z.y);
            a[3] += w * (2.f * z.x * z.y);
            a[4] += yw * z.x;
            a[5] += yw * z.y;
        }
#pragma unroll
        for (int q = 0; q < 6; ++q) acc[q] += (double)a[q];
        __syncthreads();
    }
    if (k < G.nf) {
        const CurveState st = state[ci];
        const double P = gls_power(acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], st.yy,
                                   st.ybar, pp.mode, pp.reg_c, pp.reg_s);
        if (OUT == 0) ((float*)out)[G.out + k] = (float)P;
        else ((double*)out)[G.out + k] = P;
    }
}

// The GLS is ill-conditioned where a frequency spans few cycles over the
// baseline (CC SS - CS^2 -> 0): there the FP32 round-off of the NFFT sums,
// largest at the band edge where the lowest frequencies sit, reaches |dP| ~
// 1e-4.  These few frequencies are recomputed by direct sums with FP64
// phases and FP64 accumulation.
// Grid (curve, group of 4 frequencies), the groups strided over the
// frequencies from kbegin on: each wave of a block takes one frequency and
// its 64 lanes stride over the curve's points (FP64 phases, FP64 sums), then
// reduce by shuffles.  A block stops at the first group past the curve's low
// frequencies, so however many there are (fine grids have hundreds), every
// one is covered.
constexpr int kLowPerBlock = NT / kWave;
constexpr int kLowGroups = kLowMax / kLowPerBlock;   // blocks per curve

// the six FP64 sums of frequency k over the curve's points and its power
template <int OUT, bool WIDE, int BT, int LANES>
__device__ __forceinline__ void lowfreq_one(const CurveGeom& G, const CurveState& st,
                                            const PointArrays& pa, const PlanParams& pp, int k,
                                            int tid, int lane, int N, double (*red)[6],
                                            void* __restrict__ out) {
    const double f = G.f0 + (double)k * G.df;
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = lane; j < N; j += LANES) {
        const double x = f * pa.dt[G.pts + j];
        double sn, cs;
        sincospi(2.0 * (x - rint(x)), &sn, &cs);
        const double w = pa.w[G.pts + j], yw = pa.yw[G.pts + j];
        acc[0] += w * cs;
        acc[1] += w * sn;
        acc[2] += w * (cs * cs - sn * sn);
        acc[3] += w * (2.0 * cs * sn);
        acc[4] += yw * cs;
        acc[5] += yw * sn;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = wave_sum(acc[q]);
    if (WIDE) {
        const int wv = tid / kWave;
        if ((tid & (kWave - 1)) == 0)
            for (int q = 0; q < 6; ++q) red[wv][q] = acc[q];
        __syncthreads();
        if (tid == 0)
            for (int w = 1; w < BT / kWave; ++w)
                for (int q = 0; q < 6; ++q) acc[q] += red[w][q];
        __syncthreads();   // red is reused by the block's next frequency
    }
    if (lane == 0) {
        const double P = gls_power(acc[0], acc[1], acc[2], acc[3], acc[4], acc[5], st.yy,
                                   st.ybar, pp.mode, pp.reg_c, pp.reg_s);
        if (OUT == 0) ((float*)out)[G.out + k] = (float)P;
        else ((double*)out)[G.out + k] = P;
    }
}

// WIDE: one frequency per 1024-thread block, all its lanes over the points
// (long curves: 16 waves per frequency hide the FP64 latency).
constexpr int kLowWideNT = 1024;
// LOOP: the block strides on over the grid (gridDim.y groups apart) until it
// passes the curve's low run; without it, one group per block (fewer live
// registers: the common case, launched with enough groups for kLowMax
// frequencies, and LOOP launches take any frequencies after those)
template <int OUT, bool WIDE, bool LOOP>
__global__ __launch_bounds__(WIDE ? kLowWideNT : NT) void lowfreq_kernel(
    const CurveGeom* __restrict__ geom, const int32_t* __restrict__ list,
    const CurveState* __restrict__ state, PointArrays pa, PlanParams pp, int kbegin,
    void* __restrict__ out) {
    constexpr int BT = WIDE ? kLowWideNT : NT;     // threads per block
    __shared__ double red[BT / kWave][6];
    constexpr int FPB = WIDE ? 1 : kLowPerBlock;   // frequencies per block
    constexpr int LANES = WIDE ? BT : kWave;       // lanes per frequency
    const int tid = threadIdx.x;
    const int lane = tid % LANES;
    const int c = list[blockIdx.x];
    const CurveGeom& G = geom[c];
    const CurveState& st = state[c];
    int first = kbegin + (int)blockIdx.y * FPB;
    do {
        // frequency k is low if it and every frequency before it span fewer
        // than kLowCycles cycles over the baseline (the leading run of the grid)
        if (first >= G.nf) return;
        if (fabs(G.f0 + (double)first * G.df) * st.tspan >= kLowCycles) return;
        const int k = first + tid / LANES;
        bool active = k < G.nf;
        for (int q = first; q <= k && active; ++q)
            active = fabs(G.f0 + (double)q * G.df) * st.tspan < kLowCycles;
        // uniform per frequency: whole waves (or the block) exit; later groups
        // of this block lie further along the grid, past the run as well
        if (!active) return;
        lowfreq_one<OUT, WIDE, BT, LANES>(G, st, pa, pp, k, tid, lane, G.n_obs, red, out);
        first += (int)gridDim.y * FPB;
    } while (LOOP);
}

// Wide curves with a WideAux record: grid (wide curve, frequency, part);
// each of the kLowSplit workgroups of a low frequency sums one slice of the
// points and stores its six FP64 sums in the record; the last to finish adds
// the slices in order and forms the power.  Frequencies past the low run exit at once.
template <int OUT>
__global__ __launch_bounds__(kLowWideNT) void lowfreq_split_kernel(
    const CurveGeom* __restrict__ geom, const int32_t* __restrict__ list,
    const CurveState* __restrict__ state, PointArrays pa, PlanParams pp,
    WideAux* __restrict__ aux, void* __restrict__ out) {
    __shared__ double red[kLowWideNT / kWave][6];
    const int tid = threadIdx.x;
    const int k = blockIdx.y;
    const int c = list[blockIdx.x];
    const CurveGeom& G = geom[c];
    const CurveState& st = state[c];
    if (k >= min(G.nf, kLowMax)) return;
    const double f = G.f0 + (double)k * G.df;
    if (fabs(f) * st.tspan >= kLowCycles) return;   // the test of lowfreq_kernel<OUT, true>
    const int N = G.n_obs;
    const int per = (N + kLowSplit - 1) / kLowSplit;
    const int j0 = (int)blockIdx.z * per;
    const int j1 = min(N, j0 + per);
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = j0 + tid; j < j1; j += kLowWideNT) {
        const double x = f * pa.dt[G.pts + j];
        double sn, cs;
        sincospi(2.0 * (x - rint(x)), &sn, &cs);
        const double w = pa.w[G.pts + j], yw = pa.yw[G.pts + j];
        acc[0] += w * cs;
        acc[1] += w * sn;
        acc[2] += w * (cs * cs - sn * sn);
        acc[3] += w * (2.0 * cs * sn);
        acc[4] += yw * cs;
        acc[5] += yw * sn;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] = wave_sum(acc[q]);
    const int wv = tid / kWave;
    if ((tid & (kWave - 1)) == 0)
        for (int q = 0; q < 6; ++q) red[wv][q] = acc[q];
    __syncthreads();
    if (tid != 0) return;
    for (int w = 1; w < kLowWideNT / kWave; ++w)
        for (int q = 0; q < 6; ++q) acc[q] += red[w][q];
    WideAux& X = aux[blockIdx.x];
    for (int q = 0; q < 6; ++q) X.low[k][blockIdx.z][q] = acc[q];
    __threadfence();
    if (atomicAdd(&X.low_done[k], 1) != kLowSplit - 1) return;
    __threadfence();
    // the slices in order, so the sum does not depend on which finished first
    double sm[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int z = 0; z < kLowSplit; ++z)
        for (int q = 0; q < 6; ++q)
            sm[q] += __hip_atomic_load(&X.low[k][z][q], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
    const double P = gls_power(sm[0], sm[1], sm[2], sm[3], sm[4], sm[5], st.yy, st.ybar,
                               pp.mode, pp.reg_c, pp.reg_s);
    if (OUT == 0) ((float*)out)[G.out + k] = (float)P;
    else ((double*)out)[G.out + k] = P;
}

template <int RF>
hipError_t dirsum_dispatch(const CurveGeom* d_geom, const CurveState* state,
                           const int32_t* tile_curve, const int32_t* tile_first, int ntiles,
                           PointArrays pa, const PlanParams& pp, void* power, int out_f64,
                           hipStream_t s) {
    if (out_f64)
        hipLaunchKernelGGL((dirsum_kernel<RF, 1>), dim3(ntiles), dim3(NT), 0, s, d_geom, state,
                           tile_curve, tile_first, pa, pp, power);
    else
        hipLaunchKernelGGL((dirsum_kernel<RF, 0>), dim3(ntiles), dim3(NT), 0, s, d_geom, state,
                           tile_curve, tile_first, pa, pp, power);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_lowfreq(const CurveGeom* d_geom, const int32_t* small, int nsmall,
                          const int32_t* wide, int nwide, const CurveState* state,
                          PointArrays pa, const PlanParams& pp, void* power, int out_f64,
                          WideAux* aux, hipStream_t s) {
    // the first kLowMax frequencies: one group per block; any further ones
    // (fine grids): a few looping blocks per curve, which exit at once when
    // the low run ends before kLowMax
    constexpr int kMoreGroups = 4;
#define GLS_LOW(OUTV, W, L, GRID, NTH, LIST, KB)                                                \
    hipLaunchKernelGGL((lowfreq_kernel<OUTV, W, L>), GRID, dim3(NTH), 0, s, d_geom, LIST, state, \
                       pa, pp, KB, power)
    if (nsmall > 0) {
        if (out_f64) {
            GLS_LOW(1, false, false, dim3(nsmall, kLowGroups), NT, small, 0);
            GLS_LOW(1, false, true, dim3(nsmall, kMoreGroups), NT, small, kLowMax);
        } else {
            GLS_LOW(0, false, false, dim3(nsmall, kLowGroups), NT, small, 0);
            GLS_LOW(0, false, true, dim3(nsmall, kMoreGroups), NT, small, kLowMax);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nwide > 0 && aux != nullptr) {
        // the first kLowMax frequencies split over kLowSplit workgroups each
        // (one accumulator slot per frequency in the record)
        const dim3 grid(nwide, kLowMax, kLowSplit);
        if (out_f64)
            hipLaunchKernelGGL((lowfreq_split_kernel<1>), grid, dim3(kLowWideNT), 0, s, d_geom,
                               wide, state, pa, pp, aux, power);
        else
            hipLaunchKernelGGL((lowfreq_split_kernel<0>), grid, dim3(kLowWideNT), 0, s, d_geom,
                               wide, state, pa, pp, aux, power);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    } else if (nwide > 0) {
        if (out_f64) GLS_LOW(1, true, false, dim3(nwide, kLowMax), kLowWideNT, wide, 0);
        else GLS_LOW(0, true, false, dim3(nwide, kLowMax), kLowWideNT, wide, 0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nwide > 0) {
        if (out_f64) GLS_LOW(1, true, true, dim3(nwide, kMoreGroups), kLowWideNT, wide, kLowMax);
        else GLS_LOW(0, true, true, dim3(nwide, kMoreGroups), kLowWideNT, wide, kLowMax);
    }
#undef GLS_LOW
    return hipGetLastError();
}

hipError_t launch_dirsum_custom(const CurveGeom* d_geom, const CurveState* state,
                                const int32_t* tile_curve, const int32_t* tile_first,
                                int ntiles, PointArrays pa, const PlanParams& pp,
                                const double* freqs, void* power, int out_f64, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (out_f64)
        hipLaunchKernelGGL(dirsum_custom_kernel<1>, dim3(ntiles), dim3(NT), 0, s, d_geom, state,
                           tile_curve, tile_first, pa, pp, freqs, power);
    else
        hipLaunchKernelGGL(dirsum_custom_kernel<0>, dim3(ntiles), dim3(NT), 0, s, d_geom, state,
                           tile_curve, tile_first, pa, pp, freqs, power);
    return hipGetLastError();
}

hipError_t launch_dirsum(const CurveGeom* d_geom, const CurveState* state,
                         const int32_t* tile_curve, const int32_t* tile_first, int ntiles,
                         int rf, PointArrays pa, const PlanParams& pp, void* power, int out_f64,
                         hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    switch (rf) {
        case 1: return dirsum_dispatch<1>(d_geom, state, tile_curve, tile_first, ntiles, pa, pp, power, out_f64, s);
        case 2: return dirsum_dispatch<2>(d_geom, state, tile_curve, tile_first, ntiles, pa, pp, power, out_f64, s);
        case 4: return dirsum_dispatch<4>(d_geom, state, tile_curve, tile_first, ntiles, pa, pp, power, out_f64, s);
        case 8: return dirsum_dispatch<8>(d_geom, state, tile_curve, tile_first, ntiles, pa, pp, power, out_f64, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace gls
