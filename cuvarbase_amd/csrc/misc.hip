// Gaussian gridding in the reference geometry, and per-curve top-k peaks.
#include "device_util.hpp"

namespace gls {

namespace {

constexpr int NT = 256;

// One thread per point, 2m+1 taps with global float atomics: the debug output
// of nfft_adjoint_async(just_return_gridded_data=True)
// (cuvarbase/cunfft.py:229-268; taps of cuvarbase/kernels/cunfft.cu:140-161),
// with the cell coordinate formed in FP64.
__global__ __launch_bounds__(NT) void gaussian_grid_kernel(const double* __restrict__ t,
                                                           const double* __restrict__ y,
                                                           int n_obs, int ng, int m, double b,
                                                           double tmin, double scale,
                                                           float* __restrict__ grid) {
    const int j = blockIdx.x * NT + threadIdx.x;
    if (j >= n_obs) return;
    const double x = (t[j] - tmin) * scale;
    if (!isfinite(x)) return;
    const double yc = (double)ng * (x - floor(x));
    const double fl = floor(yc);
    const int U = (int)fl - m;
    const double fr = yc - fl;
    const double yv = y[j];
    const double norm = 1.0 / sqrt(M_PI * b);
    for (int k = 0; k <= 2 * m; ++k) {
        const double d = (double)(m - k) + fr;
        int cell = U + k;
        cell %= ng;
        if (cell < 0) cell += ng;
        atomicAdd(&grid[cell], (float)(yv * exp(-d * d / b) * norm));
    }
}

// Top-k of one curve's powers (descending, ties -> lower index), skipping
// masked indices and invalid (-1 / NaN) powers: k rounds of a block argmax.
__global__ __launch_bounds__(NT) void topk_kernel(const CurveGeom* __restrict__ geom,
                                                  const float* __restrict__ power,
                                                  const uint8_t* __restrict__ mask, int K,
                                                  float* __restrict__ top_power,
                                                  int32_t* __restrict__ top_index) {
    __shared__ float sv[NT / kWave];
    __shared__ int si[NT / kWave];
    __shared__ int chosen[64];
    const int c = blockIdx.x;
    const CurveGeom& G = geom[c];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    for (int round = 0; round < K; ++round) {
        float best = -INFINITY;
        int bi = 0x7fffffff;
        for (int k = tid; k < G.nf; k += NT) {
            if (mask && mask[G.out + k]) continue;
            const float v = power[G.out + k];
            if (!(v >= 0.f)) continue;  // -1 sentinel and NaN
            bool taken = false;
            for (int q = 0; q < round; ++q) taken |= (chosen[q] == k);
            if (taken) continue;
            if (v > best || (v == best && k < bi)) {
                best = v;
                bi = k;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(best, o, kWave);
            const int oi = __shfl_xor(bi, o, kWave);
            if (ov > best || (ov == best && oi < bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            sv[wv] = best;
            si[wv] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            float bv = sv[0];
            int bx = si[0];
            for (int q = 1; q < NT / kWave; ++q) {
                if (sv[q] > bv || (sv[q] == bv && si[q] < bx)) {
                    bv = sv[q];
                    bx = si[q];
                }
            }
            const bool ok = bx != 0x7fffffff;
            chosen[round] = ok ? bx : -1;
            top_power[(int64_t)c * K + round] = ok ? bv : -1.f;
            top_index[(int64_t)c * K + round] = ok ? bx : -1;
        }
        __syncthreads();
    }
}

// Top-k for k <= KM in one read of the powers: each thread keeps a sorted
// list of its KM best (value desc, index asc) over a strided slice, then k
// rounds of block argmax over the lists' heads pick the winners in order.
// Same order and ties as topk_kernel (the order is total, so the slices and
// the order in which a thread visits its elements do not matter).  The slice
// is read UNR elements at a time, all loads issued before the first compare,
// so that one block keeps enough reads in flight for a long curve.
constexpr int NTK = 512;
template <int KM>
__global__ __launch_bounds__(NTK) void topk_small_kernel(const CurveGeom* __restrict__ geom,
                                                        const float* __restrict__ power,
                                                        const uint8_t* __restrict__ mask,
                                                        int K, float* __restrict__ top_power,
                                                        int32_t* __restrict__ top_index) {
    __shared__ float sv[NTK / kWave];
    __shared__ int si[NTK / kWave];
    __shared__ int win;
    const int c = blockIdx.x;
    const CurveGeom& G = geom[c];
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wv = tid / kWave;
    auto better = [](float v, int i, float bv, int bi) { return v > bv || (v == bv && i < bi); };
    float lv[KM];
    int li[KM];
#pragma unroll
    for (int q = 0; q < KM; ++q) {
        lv[q] = -INFINITY;
        li[q] = 0x7fffffff;
    }
    auto offer = [&](float v, int k) {
        if (!(v >= 0.f)) return;   // -1 sentinel, NaN and masked
        if (!better(v, k, lv[KM - 1], li[KM - 1])) return;
        // insert into the sorted list (registers: unrolled shift)
        float cv = v;
        int ci = k;
#pragma unroll
        for (int q = 0; q < KM; ++q) {
            if (better(cv, ci, lv[q], li[q])) {
                const float tv = lv[q];
                const int ti = li[q];
                lv[q] = cv;
                li[q] = ci;
                cv = tv;
                ci = ti;
            }
        }
    };
    constexpr int UNR = 4;
    const float* const pw = power + G.out;
    const uint8_t* const mk = mask ? mask + G.out : nullptr;
    const int nf = G.nf;
    int k = tid;
    for (; k + (UNR - 1) * NTK < nf; k += UNR * NTK) {
        float v[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) v[u] = pw[k + u * NTK];
        if (mk) {
#pragma unroll
            for (int u = 0; u < UNR; ++u)
                if (mk[k + u * NTK]) v[u] = -1.f;
        }
#pragma unroll
        for (int u = 0; u < UNR; ++u) offer(v[u], k + u * NTK);
    }
    for (; k < nf; k += NTK) offer((mk && mk[k]) ? -1.f : pw[k], k);
    int head = 0;   // next unpicked entry of this thread's list
    for (int round = 0; round < K; ++round) {
        float best = -INFINITY;
        int bi = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < KM; ++q)
            if (q == head) {
                best = lv[q];
                bi = li[q];
            }
        const int mine = bi;   // this thread's head (indices are unique per curve)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const float ov = __shfl_xor(best, o, kWave);
            const int oi = __shfl_xor(bi, o, kWave);
            if (better(ov, oi, best, bi)) {
                best = ov;
                bi = oi;
            }
        }
        if (lane == 0) {
            sv[wv] = best;
            si[wv] = bi;
        }
        __syncthreads();
        if (tid == 0) {
            float bv = sv[0];
            int bx = si[0];
            for (int q = 1; q < NTK / kWave; ++q)
                if (better(sv[q], si[q], bv, bx)) {
                    bv = sv[q];
                    bx = si[q];
                }
            const bool ok = bx != 0x7fffffff;
            top_power[(int64_t)c * K + round] = ok ? bv : -1.f;
            top_index[(int64_t)c * K + round] = ok ? bx : -1;
            win = bx;
        }
        __syncthreads();
        // the thread whose head won moves on to its next entry
        if (win != 0x7fffffff && mine == win) ++head;
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_gaussian_grid(const double* t, const double* y, int n_obs, int ng, int m,
                                double b, double tmin, double tmax, double spp, float* grid,
                                hipStream_t s) {
    hipError_t e = hipMemsetAsync(grid, 0, sizeof(float) * (size_t)ng, s);
    if (e != hipSuccess) return e;
    if (n_obs <= 0) return hipSuccess;
    const double scale = 1.0 / (spp * (tmax - tmin));
    hipLaunchKernelGGL(gaussian_grid_kernel, dim3((n_obs + NT - 1) / NT), dim3(NT), 0, s, t, y,
                       n_obs, ng, m, b, tmin, scale, grid);
    return hipGetLastError();
}

hipError_t launch_topk(const CurveGeom* d_geom, int ncurves, const float* power,
                       const uint8_t* mask, int k, float* top_power, int32_t* top_index,
                       hipStream_t s) {
    if (ncurves <= 0 || k <= 0) return hipSuccess;
    if (k <= 8)
        hipLaunchKernelGGL((topk_small_kernel<8>), dim3(ncurves), dim3(NTK), 0, s, d_geom, power,
                           mask, k, top_power, top_index);
    else
        hipLaunchKernelGGL(topk_kernel, dim3(ncurves), dim3(NT), 0, s, d_geom, power, mask, k,
                           top_power, top_index);
    return hipGetLastError();
}

}  // namespace gls
