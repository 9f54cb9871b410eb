// Small device helpers: wave64/block reductions, phase reduction.
#pragma once

#include "gls_internal.hpp"

namespace gls {

constexpr int kWave = 64;
constexpr int kNumXcd = 8;   // MI355X: 8 accelerator complex dies, one L2 each

// The dispatcher hands consecutive workgroups to the XCDs in turn.  This maps
// the hardware index to a logical one so that each XCD runs a contiguous range
// of logical blocks: neighbouring tiles (which share cache lines of the
// intermediate grid and the same point lists) then meet in the same L2.
// A bijection of [0, nblocks) for any nblocks.
__host__ __device__ constexpr int xcd_logical_block(int bid, int nblocks) {
    const int q = nblocks / kNumXcd;
    const int r = nblocks % kNumXcd;
    const int x = bid % kNumXcd;
    const int i = bid / kNumXcd;
    return (x < r) ? x * (q + 1) + i : r * (q + 1) + (x - r) * q + i;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, kWave);
        v = u < v ? u : v;
    }
    return v;
}
template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T u = __shfl_xor(v, o, kWave);
        v = u > v ? u : v;
    }
    return v;
}

// Block-wide reductions for blockDim.x == NT (multiple of 64).  `scratch`
// must hold NT/64 values; the result is returned to every thread.
template <int NT, typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) scratch[w] = v;
    __syncthreads();
    T r = 0;
#pragma unroll
    for (int i = 0; i < NT / kWave; ++i) r += scratch[i];
    __syncthreads();
    return r;
}
template <int NT, typename T>
__device__ __forceinline__ T block_min(T v, T* scratch) {
    v = wave_min(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) scratch[w] = v;
    __syncthreads();
    T r = scratch[0];
#pragma unroll
    for (int i = 1; i < NT / kWave; ++i) r = scratch[i] < r ? scratch[i] : r;
    __syncthreads();
    return r;
}
template <int NT, typename T>
__device__ __forceinline__ T block_max(T v, T* scratch) {
    v = wave_max(v);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) scratch[w] = v;
    __syncthreads();
    T r = scratch[0];
#pragma unroll
    for (int i = 1; i < NT / kWave; ++i) r = scratch[i] > r ? scratch[i] : r;
    __syncthreads();
    return r;
}

// Exponent e of the fixed-point spreading scale 2^e of an LDS tile whose
// cells can hold at most `bound` in magnitude: scaled cells stay below 2^30,
// half the int32 range.  Clamped to the FP32 exponent range.
__device__ __forceinline__ int spread_exponent(float bound) {
    if (!(bound > 0.0f) || !isfinite(bound)) return 0;
    int k;
    (void)frexpf(bound, &k);   // bound < 2^k
    const int e = 30 - k;
    return e < -120 ? -120 : (e > 120 ? 120 : e);
}

// exp(2 pi i r) for a phase r given in revolutions as FP64: the integer part
// is removed in FP64 (r - rint(r) in [-1/2, 1/2]) and only the reduced phase
// is evaluated in FP32.
__device__ __forceinline__ float2 expi_rev(double r) {
    const float rr = (float)(r - rint(r));
    float s, c;
    sincospif(2.0f * rr, &s, &c);
    return make_float2(c, s);
}

// Buffer access to one curve's grid region (at most 3 grids of 2^24
// complex64 cells, under 2^32 bytes): the base address and size live in a
// scalar resource, every access takes a 32-bit byte offset, so the per-access
// address arithmetic is one 32-bit add rather than a 64-bit multiply-add, and
// offsets past the region read zero / drop the store.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t grid_rsrc(const void* base, uint32_t bytes) {
    // 4th descriptor word: raw 32-bit access (the value gfx9 raw buffer
    // accesses use)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                             0x00020000);
}
typedef float gls_v2f __attribute__((ext_vector_type(2)));
typedef unsigned int gls_v2u __attribute__((ext_vector_type(2)));
// byte offset voff (per lane) + soff (wave-uniform: a scalar register, no
// vector arithmetic); callers keep voff + soff inside the region
__device__ __forceinline__ float2 grid_load(__amdgpu_buffer_rsrc_t r, uint32_t voff,
                                           uint32_t soff = 0) {
    const gls_v2f v =
        __builtin_bit_cast(gls_v2f, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
    return make_float2(v.x, v.y);
}
// CPOL: cache policy of the store, 0 the default, kStoreNonTemporal streaming
// (the line is not kept in the caches)
constexpr int kStoreNonTemporal = 2;
template <int CPOL = 0>
__device__ __forceinline__ void grid_store(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff,
                                           float2 v) {
    const gls_v2f x = {v.x, v.y};
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(gls_v2u, x), r, voff, soff, CPOL);
}
__device__ __forceinline__ void buf_store_f32(__amdgpu_buffer_rsrc_t r, uint32_t voff, float v) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, voff, 0u, 0);
}

// GLS power from the six sums, closed form b^T A^-1 b / den in FP64
// (algebraically equal to lspow / lspow_flmean / lspow0 of
// cuvarbase/kernels/lomb.cu:36-141; invalid -> -1 as lomb.cu:79-80).
__device__ __forceinline__ double gls_power(double C, double S, double C2, double S2, double YC,
                                            double YS, double YY, double ybar, int mode,
                                            double l0, double l1) {
    double CC, SS, CS, bc, bs, den;
    if (mode == GLS_MODE_FLOATING_MEAN) {
        CC = 0.5 * (1.0 + C2) - C * C + l0;
        SS = 0.5 * (1.0 - C2) - S * S + l1;
        CS = 0.5 * S2 - C * S;
        bc = YC;
        bs = YS;
        den = YY;
    } else if (mode == GLS_MODE_STANDARD) {
        CC = 0.5 * (1.0 + C2) + l0;
        SS = 0.5 * (1.0 - C2) + l1;
        CS = 0.5 * S2;
        bc = YC + ybar * C;
        bs = YS + ybar * S;
        den = YY + ybar * ybar;
    } else if (mode == GLS_MODE_WINDOW) {
        CC = 0.5 * (1.0 + C2);
        SS = 0.5 * (1.0 - C2);
        CS = 0.5 * S2;
        bc = 2.0 * C;
        bs = 2.0 * S;
        den = 1.0;
    } else {
        return -1.0;
    }
    const double D = CC * SS - CS * CS;
    const double P = (SS * bc * bc + CC * bs * bs - 2.0 * CS * bc * bs) / (den * D);
    if (!isfinite(P) || P < 0.0) return -1.0;
    return P;
}

// The same power with the sums and their combinations in FP32, as the NFFT
// assembly runs: the determinant CC SS - CS^2 and the numerator are formed in
// FP64 from the FP32 terms (the products are exact there, so their
// cancellation costs no digits beyond those already in CC, SS, CS), and the
// quotient in FP32.  Frequencies with f (t_max - t_min) < 3, where CC, SS and
// CS cancel themselves, are recomputed by the FP64 direct sums.  den_std is
// the standard mode's denominator YY + ybar^2, formed once by the caller.
__device__ __forceinline__ float gls_power_f32(float C, float S, float C2, float S2, float YC,
                                               float YS, float YY, float ybar, int mode,
                                               float l0, float l1, float den_std) {
    float CC, SS, CS, bc, bs, den;
    if (mode == GLS_MODE_FLOATING_MEAN) {
        CC = 0.5f * (1.0f + C2) - C * C + l0;
        SS = 0.5f * (1.0f - C2) - S * S + l1;
        CS = 0.5f * S2 - C * S;
        bc = YC;
        bs = YS;
        den = YY;
    } else if (mode == GLS_MODE_STANDARD) {
        CC = 0.5f * (1.0f + C2) + l0;
        SS = 0.5f * (1.0f - C2) + l1;
        CS = 0.5f * S2;
        bc = YC + ybar * C;
        bs = YS + ybar * S;
        den = den_std;
    } else if (mode == GLS_MODE_WINDOW) {
        CC = 0.5f * (1.0f + C2);
        SS = 0.5f * (1.0f - C2);
        CS = 0.5f * S2;
        bc = 2.0f * C;
        bs = 2.0f * S;
        den = 1.0f;
    } else {
        return -1.0f;
    }
    const double cc = CC, ss = SS, cs = CS, xc = bc, xs = bs;
    const double D = cc * ss - cs * cs;
    // N = b^T adj(A) b = SS bc^2 + CC bs^2 - 2 CS bc bs
    const double N = xc * (ss * xc - cs * xs) + xs * (cc * xs - cs * xc);
    const float P = (float)N * __builtin_amdgcn_rcpf(den * (float)D);
    if (!isfinite(P) || P < 0.0f) return -1.0f;
    return P;
}

}  // namespace gls
