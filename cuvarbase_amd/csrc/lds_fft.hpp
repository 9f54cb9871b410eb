// LDS-resident mixed-radix Stockham FFT building blocks for gfx950 (wave64).
//
// All transforms use the + sign of an unnormalised inverse DFT,
//   X[k] = sum_u x[u] exp(+2 pi i u k / L),
// the convention the reference obtains from cuFFT's inverse C2C
// (cuvarbase/cunfft.py:288, pinned by cuvarbase/tests/test_nfft.py:186-200).
// Lengths are 2^a 3^b 5^c <= 4096; stages use radices 16, 12, 9, 8, 4, 2, 5, 3.
// Stage twiddles come from a two-level table of exp(2 pi i x / L) in LDS
// (64 + ceil(L/64) entries) and integer powers by complex products.
#pragma once

#include "gls_internal.hpp"

namespace gls {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 mul_i(float2 a) { return make_float2(-a.y, a.x); }

constexpr int ilog2_c(int v) { return v <= 1 ? 0 : 1 + ilog2_c(v >> 1); }
constexpr int brev_c(int v, int bits) {
    return bits == 0 ? 0 : ((v & 1) << (bits - 1)) | brev_c(v >> 1, bits - 1);
}
// radix of the next Stockham stage for a remaining length rem: the largest
// in-register DFT that divides it (fewer stages = fewer LDS round trips)
constexpr int next_radix(int rem) {
    return rem % 16 == 0 ? 16 : rem % 12 == 0 ? 12 : rem % 9 == 0 ? 9 : rem % 8 == 0 ? 8
         : rem % 4 == 0 ? 4 : rem % 2 == 0 ? 2 : rem % 5 == 0 ? 5 : rem % 3 == 0 ? 3 : rem;
}

// compile-time cos / sin (Taylor series after reduction to [-pi, pi]) for the
// twiddles inside composite in-register DFTs
constexpr double kPiC = 3.14159265358979323846;
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 24; ++i) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int i = 1; i < 24; ++i) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
// exp(+2 pi i m / R), m = 0 .. R-1
template <int R>
struct RootsOfUnity {
    float c[R];
    float s[R];
    constexpr RootsOfUnity() : c(), s() {
        for (int m = 0; m < R; ++m) {
            double x = 2.0 * kPiC * m / R;
            if (x > kPiC) x -= 2.0 * kPiC;
            c[m] = (float)taylor_cos(x);
            s[m] = (float)taylor_sin(x);
        }
    }
};
constexpr bool smooth235(int v) {
    while (v % 2 == 0) v /= 2;
    while (v % 3 == 0) v /= 3;
    while (v % 5 == 0) v /= 5;
    return v == 1;
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

template <int R>
__device__ __forceinline__ void dft_reg(float2 (&x)[R]);

// composite in-register DFT of length R1 R2 (Cooley-Tukey with n = n1 + R1 n2,
// k = R2 k1 + k2): length-R2 DFTs over n2, twiddles exp(2 pi i n1 k2 / R),
// length-R1 DFTs over n1
template <int R1, int R2>
__device__ __forceinline__ void dft_ct(float2 (&x)[R1 * R2]) {
    constexpr int R = R1 * R2;
    constexpr RootsOfUnity<R> W{};
    float2 y[R1][R2];
#pragma unroll
    for (int n1 = 0; n1 < R1; ++n1) {
#pragma unroll
        for (int n2 = 0; n2 < R2; ++n2) y[n1][n2] = x[n1 + R1 * n2];
        dft_reg<R2>(y[n1]);
#pragma unroll
        for (int k2 = 1; k2 < R2; ++k2) {
            const int e = (n1 * k2) % R;
            if (e != 0) y[n1][k2] = cmul(y[n1][k2], make_float2(W.c[e], W.s[e]));
        }
    }
#pragma unroll
    for (int k2 = 0; k2 < R2; ++k2) {
        float2 z[R1];
#pragma unroll
        for (int n1 = 0; n1 < R1; ++n1) z[n1] = y[n1][k2];
        dft_reg<R1>(z);
#pragma unroll
        for (int k1 = 0; k1 < R1; ++k1) x[R2 * k1 + k2] = z[k1];
    }
}

// In-register DFT: x[r] <- sum_q x[q] exp(+2 pi i q r / R)
template <int R>
__device__ __forceinline__ void dft_reg(float2 (&x)[R]) {
    if constexpr (R == 9) {
        dft_ct<3, 3>(x);
    } else if constexpr (R == 12) {
        dft_ct<4, 3>(x);
    } else if constexpr (R == 3) {
        constexpr float h = 0.86602540378443864676f;   // sqrt(3)/2
        const float2 t1 = cadd(x[1], x[2]);
        const float2 t2 = csub(x[1], x[2]);
        const float2 m1 = make_float2(x[0].x - 0.5f * t1.x, x[0].y - 0.5f * t1.y);
        const float2 m2 = make_float2(-h * t2.y, h * t2.x);
        x[0] = cadd(x[0], t1);
        x[1] = cadd(m1, m2);
        x[2] = csub(m1, m2);
    } else if constexpr (R == 5) {
        constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi/5)
        constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi/5)
        constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi/5)
        constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi/5)
        const float2 t1 = cadd(x[1], x[4]);
        const float2 t2 = cadd(x[2], x[3]);
        const float2 t3 = csub(x[1], x[4]);
        const float2 t4 = csub(x[2], x[3]);
        const float2 a1 = make_float2(x[0].x + c1 * t1.x + c2 * t2.x, x[0].y + c1 * t1.y + c2 * t2.y);
        const float2 a2 = make_float2(x[0].x + c2 * t1.x + c1 * t2.x, x[0].y + c2 * t1.y + c1 * t2.y);
        const float2 b1 = mul_i(make_float2(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y));
        const float2 b2 = mul_i(make_float2(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y));
        x[0] = cadd(x[0], cadd(t1, t2));
        x[1] = cadd(a1, b1);
        x[4] = csub(a1, b1);
        x[2] = cadd(a2, b2);
        x[3] = csub(a2, b2);
    } else {
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
}

// exp(2 pi i q / n) for 0 <= q < 2^31 and 0 < n <= 2^24, evaluated from the
// exactly reduced ratio (|2q'| <= n is exact in FP32) with a 32-bit remainder
__device__ __forceinline__ float2 expi_ratio32(uint32_t q, uint32_t n) {
    const uint32_t r = q % n;
    const int rs = (2 * r >= n) ? (int)r - (int)n : (int)r;
    float s, c;
    sincospif((float)(2 * rs) / (float)n, &s, &c);
    return make_float2(c, s);
}

// two-level table of exp(2 pi i x / L), x < L: hi[x >> 6] * lo[x & 63]
template <int L>
struct TwTable {
    static constexpr int NH = (L + 63) / 64;
    float2 hi[NH];
    float2 lo[64];
};

// The same table as compile-time constants (cos, sin of 2 pi x / L for the
// lo entries x = 0..63, then the hi entries x = 64 h), evaluated by the
// compiler in double precision and rounded once to float; blocks copy it into
// LDS instead of evaluating a sine and cosine per entry.
template <int L>
struct TwConst {
    static constexpr int N = 64 + TwTable<L>::NH;
    float c[N];
    float s[N];
    constexpr TwConst() : c(), s() {
        for (int i = 0; i < N; ++i) {
            long long x = (i < 64 ? (long long)i : (long long)(i - 64) * 64) % L;
            if (2 * x >= L) x -= L;
            const double a = 2.0 * kPiC * (double)x / L;
            c[i] = (float)taylor_cos(a);
            s[i] = (float)taylor_sin(a);
        }
    }
};
template <int L>
__device__ constexpr TwConst<L> kTwConst{};

template <int L, int NT>
__device__ __forceinline__ void build_twiddles(TwTable<L>& T, int tid) {
    const TwConst<L>& C = kTwConst<L>;
    for (int i = tid; i < TwConst<L>::N; i += NT) {
        const float2 v = make_float2(C.c[i], C.s[i]);
        if (i < 64) T.lo[i] = v;
        else T.hi[i - 64] = v;
    }
}

template <int L>
__device__ __forceinline__ float2 tw_lookup(const TwTable<L>& T, int x) {
    if (L <= 64) return T.lo[x];
    const float2 h = T.hi[x >> 6];
    const int l = x & 63;
    return l == 0 ? h : cmul(h, T.lo[l]);
}

// LDS placement of element i of interleaved sequence b: flattened index
// f = i*BS + b, plus one pad slot after every 16 elements.  The pad breaks the
// power-of-two strides of the Stockham writes ((j - k) R + k + q NS), which
// would otherwise put most lanes of a wave on the same LDS banks.
template <int BS>
__device__ __forceinline__ int lds_at(int i, int b) {
    const int f = i * BS + b;
    return f + (f >> 4);
}
// LDS elements needed for L rows of stride BS (largest lds_at index + 1)
constexpr int lds_size(int L, int BS) { return L * BS + ((L * BS) >> 4) + 1; }

// Element i0 + q*step of sequence b, q = 0, 1, ...: when step*BS is a multiple
// of 16 the pad slots grow linearly with q, so the places are
// lds_at(i0, b) + q*(step*BS*17/16) -- compile-time offsets from one base,
// which the LDS instructions take as immediates instead of a shift, mask and
// add per element.
template <int BS, int STEP>
struct LdsRun {
    static constexpr bool LINEAR = (STEP * BS) % 16 == 0;
    static constexpr int DELTA = STEP * BS + (STEP * BS) / 16;
    int base, i0, b;
    __device__ __forceinline__ LdsRun(int i0_, int b_) : base(lds_at<BS>(i0_, b_)), i0(i0_), b(b_) {}
    __device__ __forceinline__ int at(int q) const {
        if constexpr (LINEAR) return base + q * DELTA;
        else return lds_at<BS>(i0 + q * STEP, b);
    }
};

// One radix-R Stockham stage over B interleaved sequences of length L held in
// LDS at s[lds_at<BS>(i, b)].  NS = product of the radices of earlier stages.
// Butterfly beta (b = beta % B, j = beta / B) reads x[j + q L/R], multiplies
// input q by w^q, w = exp(2 pi i k / (NS R)), k = j % NS, and its output q
// goes to index (j - k) R + k + q NS.
template <int L, int B, int BS, int R, int NS, int NT>
struct StockhamStage {
    static constexpr int LR = L / R;
    static constexpr int NBUT = LR * B;               // butterflies in the stage
    static constexpr int NB = (NBUT + NT - 1) / NT;   // per thread

    static __device__ __forceinline__ bool active(int beta) {
        return NBUT % NT == 0 || beta < NBUT;
    }

    // read + twiddle + radix-R DFT of this thread's butterflies
    static __device__ __forceinline__ void compute(const float2* s, const TwTable<L>& T,
                                                   int tid, float2 (&x)[NB][R]) {
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int beta = tid + r * NT;
            if (active(beta)) {
                const int b = beta % B;
                const int j = beta / B;
                {
                    const LdsRun<BS, LR> run(j, b);
#pragma unroll
                    for (int q = 0; q < R; ++q) x[r][q] = s[run.at(q)];
                }
                if constexpr (NS > 1) {
                    const int k = j % NS;
                    float2 w[R];
                    w[1] = tw_lookup<L>(T, k * (L / (NS * R)));
#pragma unroll
                    for (int q = 2; q < R; ++q) {
                        const int h = (q & (q - 1)) ? (1 << ilog2_c(q)) : (q >> 1);
                        w[q] = cmul(w[h], w[q - h]);
                    }
#pragma unroll
                    for (int q = 1; q < R; ++q) x[r][q] = cmul(x[r][q], w[q]);
                }
                dft_reg<R>(x[r]);
            }
        }
    }

    // first stage reading its inputs from src(j, q, LR, b) -- element j + q LR of
    // sequence b -- instead of LDS
    template <class S>
    static __device__ __forceinline__ void compute_from(S&& src, int tid,
                                                        float2 (&x)[NB][R]) {
        static_assert(NS == 1, "only the first stage reads from a source");
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int beta = tid + r * NT;
            if (active(beta)) {
                const int b = beta % B;
                const int j = beta / B;
#pragma unroll
                for (int q = 0; q < R; ++q) x[r][q] = src(j, q, LR, b);
            }
        }
#pragma unroll
        for (int r = 0; r < NB; ++r)
            if (active(tid + r * NT)) dft_reg<R>(x[r]);
    }

    // hand every output to f(base, q, NS, b, v): output q of a butterfly has
    // index base + q NS in sequence b
    template <class F>
    static __device__ __forceinline__ void emit_split(int tid, const float2 (&x)[NB][R], F&& f) {
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int beta = tid + r * NT;
            if (active(beta)) {
                const int b = beta % B;
                const int j = beta / B;
                const int k = j % NS;
                const int base = (j - k) * R + k;
#pragma unroll
                for (int q = 0; q < R; ++q) f(base, q, NS, b, x[r][q]);
            }
        }
    }

    // the outputs back into LDS (output q of a butterfly at base + q NS)
    static __device__ __forceinline__ void emit_lds(float2* s, int tid, const float2 (&x)[NB][R]) {
#pragma unroll
        for (int r = 0; r < NB; ++r) {
            const int beta = tid + r * NT;
            if (active(beta)) {
                const int b = beta % B;
                const int j = beta / B;
                const int k = j % NS;
                const LdsRun<BS, NS> run((j - k) * R + k, b);
#pragma unroll
                for (int q = 0; q < R; ++q) s[run.at(q)] = x[r][q];
            }
        }
    }

    // in-place stage: returns synchronised
    static __device__ __forceinline__ void run(float2* s, const TwTable<L>& T, int tid) {
        float2 x[NB][R];
        compute(s, T, tid, x);
        __syncthreads();
        emit_lds(s, tid, x);
        __syncthreads();
    }
};

template <int L, int B, int BS, int NS, int NT>
struct LdsFftStages {
    static __device__ __forceinline__ void run(float2* s, const TwTable<L>& T, int tid) {
        if constexpr (NS < L) {
            constexpr int R = next_radix(L / NS);
            StockhamStage<L, B, BS, R, NS, NT>::run(s, T, tid);
            LdsFftStages<L, B, BS, NS * R, NT>::run(s, T, tid);
        }
    }
    // all stages in place except the last, whose outputs go to
    // sink(base, q, NS, b, v): index base + q NS of sequence b
    // straight from registers (no LDS write, no barrier after it)
    template <class F>
    static __device__ __forceinline__ void run_to(float2* s, const TwTable<L>& T, int tid,
                                                  F&& sink) {
        constexpr int R = next_radix(L / NS);
        using St = StockhamStage<L, B, BS, R, NS, NT>;
        if constexpr (NS * R < L) {
            St::run(s, T, tid);
            LdsFftStages<L, B, BS, NS * R, NT>::run_to(s, T, tid, sink);
        } else {
            float2 x[St::NB][R];
            St::compute(s, T, tid, x);
            St::emit_split(tid, x, sink);
        }
    }
};

// Full in-place FFT (natural order in and out) of B sequences of length L
// stored at s[lds_at<BS>(i, b)].  The caller has synchronised after writing s and T;
// returns synchronised.
template <int L, int B, int BS, int NT>
__device__ __forceinline__ void lds_fft(float2* s, const TwTable<L>& T, int tid) {
    static_assert(L >= 2 && L <= 4096 && smooth235(L), "L must be 2^a 3^b 5^c <= 4096");
    LdsFftStages<L, B, BS, 1, NT>::run(s, T, tid);
}

// The same transform with the first stage's inputs read from src(j, q, LR, b)
// (element j + q LR of sequence b, LR = L / first radix)
// (natural order index i of sequence b) rather than from s: the loads of all
// of a thread's first-stage butterflies are in flight together.  s need not
// be initialised; T must be complete at the first barrier.  Returns
// synchronised with the result in s.
template <int L, int B, int BS, int NT, class S>
__device__ __forceinline__ void lds_fft_from(float2* s, const TwTable<L>& T, int tid, S&& src) {
    static_assert(L >= 2 && L <= 4096 && smooth235(L), "L must be 2^a 3^b 5^c <= 4096");
    constexpr int R = next_radix(L);
    using St = StockhamStage<L, B, BS, R, 1, NT>;
    float2 x[St::NB][R];
    St::compute_from(src, tid, x);
    St::emit_lds(s, tid, x);
    __syncthreads();
    LdsFftStages<L, B, BS, R, NT>::run(s, T, tid);
}

// The same transform with the last stage's outputs handed to
// sink(base, q, NS, b, v) (index base + q NS of sequence b)
// (natural order index i of sequence b) instead of being stored; s is left
// holding the second-to-last stage.  No barrier at the end.
template <int L, int B, int BS, int NT, class F>
__device__ __forceinline__ void lds_fft_to(float2* s, const TwTable<L>& T, int tid, F&& sink) {
    static_assert(L >= 2 && L <= 4096 && smooth235(L), "L must be 2^a 3^b 5^c <= 4096");
    LdsFftStages<L, B, BS, 1, NT>::run_to(s, T, tid, sink);
}

// LDS row stride for B interleaved sequences (bank spreading comes from the
// pad slots of lds_at)
constexpr int lds_stride(int B) { return B; }

}  // namespace gls
