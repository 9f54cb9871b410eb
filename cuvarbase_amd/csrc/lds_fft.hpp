// LDS-resident Stockham FFT building blocks for gfx950 (wave64).
//
// All transforms use the + sign of an unnormalised inverse DFT,
//   X[k] = sum_u x[u] exp(+2 pi i u k / L),
// the convention the reference obtains from cuFFT's inverse C2C
// (cuvarbase/cunfft.py:288, pinned by cuvarbase/tests/test_nfft.py:186-200).
#pragma once

#include "gls_internal.hpp"

namespace gls {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v >> 1); }
constexpr int brev_c(int v, int bits) {
    return bits == 0 ? 0 : ((v & 1) << (bits - 1)) | brev_c(v >> 1, bits - 1);
}

// v * exp(+2 pi i e / 16) with e known after unrolling (constants fold)
__device__ __forceinline__ float2 mul_w16(float2 v, int e) {
    constexpr float c1 = 0.92387953251128675613f;   // cos(pi/8)
    constexpr float s1 = 0.38268343236508977173f;   // sin(pi/8)
    constexpr float r2 = 0.70710678118654752440f;   // sqrt(1/2)
    switch (e & 15) {
        case 0: return v;
        case 1: return cmul(v, make_float2(c1, s1));
        case 2: return make_float2(r2 * (v.x - v.y), r2 * (v.x + v.y));
        case 3: return cmul(v, make_float2(s1, c1));
        case 4: return make_float2(-v.y, v.x);
        case 5: return cmul(v, make_float2(-s1, c1));
        case 6: return make_float2(-r2 * (v.x + v.y), r2 * (v.x - v.y));
        case 7: return cmul(v, make_float2(-c1, s1));
        case 8: return make_float2(-v.x, -v.y);
        case 9: return cmul(v, make_float2(-c1, -s1));
        case 10: return make_float2(r2 * (v.y - v.x), -r2 * (v.x + v.y));
        case 11: return cmul(v, make_float2(-s1, -c1));
        case 12: return make_float2(v.y, -v.x);
        case 13: return cmul(v, make_float2(s1, -c1));
        case 14: return make_float2(r2 * (v.x + v.y), r2 * (v.y - v.x));
        default: return cmul(v, make_float2(c1, -s1));
    }
}

// In-register DFT of size R in {2,4,8,16}: x[r] <- sum_q x[q] exp(+2 pi i q r / R)
template <int R>
__device__ __forceinline__ void dft_reg(float2 (&x)[R]) {
    constexpr int LOG = ilog2_c(R);
    float2 y[R];
#pragma unroll
    for (int i = 0; i < R; ++i) y[brev_c(i, LOG)] = x[i];
#pragma unroll
    for (int len = 2; len <= R; len <<= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int i = 0; i < R; i += len) {
#pragma unroll
            for (int j = 0; j < half; ++j) {
                float2 u = y[i + j];
                float2 v = mul_w16(y[i + j + half], j * (16 / len));
                y[i + j] = cadd(u, v);
                y[i + j + half] = csub(u, v);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < R; ++i) x[i] = y[i];
}

// exp(2 pi i q / 2^p), q < 2^p, p <= 24, from the resident two-level table
__device__ __forceinline__ float2 twiddle_pow2(const Twiddles& tw, uint32_t q, int p) {
    uint32_t Q = q << (24 - p);
    float2 h = tw.hi[Q >> 12];
    uint32_t lo = Q & 4095u;
    return lo == 0 ? h : cmul(h, tw.lo[lo]);
}

// One radix-R Stockham stage over B interleaved sequences of length L held in
// LDS as s[i*BS + b].  NS = product of the radices of earlier stages.
template <int L, int B, int BS, int R, int NS, int NT>
__device__ __forceinline__ void stockham_stage(float2* s, const Twiddles& tw, int tid) {
    constexpr int NBUT = (L / R) * B;                 // butterflies in the stage
    constexpr int NB = (NBUT + NT - 1) / NT;          // per thread
    constexpr int LOGL = ilog2_c(L);
    float2 x[NB][R];
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int beta = tid + r * NT;
        if (NBUT % NT == 0 || beta < NBUT) {
            const int b = beta % B;
            const int j = beta / B;
#pragma unroll
            for (int q = 0; q < R; ++q) x[r][q] = s[(j + q * (L / R)) * BS + b];
            if (NS > 1) {
                const int k = j & (NS - 1);
#pragma unroll
                for (int q = 1; q < R; ++q) {
                    // exp(2 pi i q k / (NS R)) = hi[(q k L/(NS R)) << (12 - LOGL)]
                    const int e = (q * k * (L / (NS * R))) << (12 - LOGL);
                    x[r][q] = cmul(x[r][q], tw.hi[e]);
                }
            }
            dft_reg<R>(x[r]);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < NB; ++r) {
        const int beta = tid + r * NT;
        if (NBUT % NT == 0 || beta < NBUT) {
            const int b = beta % B;
            const int j = beta / B;
            const int k = j & (NS - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int q = 0; q < R; ++q) s[(base + q * NS) * BS + b] = x[r][q];
        }
    }
    __syncthreads();
}

template <int L, int B, int BS, int NS, int NT>
struct LdsFftStages {
    static __device__ __forceinline__ void run(float2* s, const Twiddles& tw, int tid) {
        if constexpr (NS < L) {
            constexpr int REM = L / NS;
            constexpr int R = REM >= 16 ? 16 : REM;
            stockham_stage<L, B, BS, R, NS, NT>(s, tw, tid);
            LdsFftStages<L, B, BS, NS * R, NT>::run(s, tw, tid);
        }
    }
};

// Full in-place FFT (natural order in, natural order out) of B sequences of
// length L (power of two, 2 <= L <= 4096) stored at s[i*BS + b].
// Caller must __syncthreads() before (data written) -- returns synchronised.
template <int L, int B, int BS, int NT>
__device__ __forceinline__ void lds_fft(float2* s, const Twiddles& tw, int tid) {
    static_assert(L >= 2 && L <= 4096 && (L & (L - 1)) == 0, "L must be 2^k <= 4096");
    LdsFftStages<L, B, BS, 1, NT>::run(s, tw, tid);
}

// padded LDS row stride for B interleaved sequences (odd stride avoids bank
// conflicts when lanes walk the sequence index)
constexpr int lds_stride(int B) { return B >= 4 ? B + 1 : B; }

}  // namespace gls
