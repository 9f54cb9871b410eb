// Per-curve data preparation and point bucketing.
//
// One workgroup per curve (1024 threads for curves of more than 16,384
// points, 256 otherwise):
//   phase A  reductions: tmin/tmax, and for raw input the weighted mean and
//            variance of LombScargleMemory.setdata (cuvarbase/lombscargle.py:
//            243-257) after the nanmean centring of normalize_light_curves
//            (cuvarbase/utils.py:134-141);
//   phase B  (NFFT method) for each of the two node sets -- x = df (t - tref)
//            for SYW and SW(f), 2x for SW(2f) -- the grid cell U, the in-cell
//            offset and the band-centring modulation exp(2 pi i Fc (t - tref))
//            of every point, in FP64 before the cast to FP32, then a counting
//            sort of the points by pass-1 column block (so every pass-1 tile
//            reads only the points that touch it); for curves of more than
//            16,384 points this phase runs as a counting kernel and a scatter
//            kernel over chunks of 2,048 points (LDS counts per chunk, one
//            global atomic per bucket and chunk);
//            (direct method) the FP32 weights and FP64 time offsets.
// With FP32 coordinates (GLS_PRECISION_FP32, for comparison with the
// reference's arithmetic) t is rounded to FP32 after centring on its mean and
// the time offset, cell coordinate and modulation phase of every point are
// FP32 operations; everything else is unchanged.
#include "device_util.hpp"

namespace gls {

namespace {

constexpr int NT = 256;

struct PrepVals {
    double t;
    double yv;   // raw: y; weighted: yw; nfft: y
    double wv;   // raw: dy; weighted: w
};

template <int KIND>  // 0 raw, 1 weighted, 2 nfft
__device__ __forceinline__ PrepVals load_point(const double* t, const void* a, const void* b,
                                               int64_t i, int y_is_f32) {
    PrepVals p;
    p.t = t[i];
    if (KIND == 0) {
        p.yv = ((const double*)a)[i];
        p.wv = ((const double*)b)[i];
    } else if (KIND == 1) {
        p.yv = (double)((const float*)a)[i];
        p.wv = (double)((const float*)b)[i];
    } else {
        p.yv = y_is_f32 ? (double)((const float*)a)[i] : ((const double*)a)[i];
        p.wv = 1.0;
    }
    return p;
}

// time offset of a point from the curve's origin: t - tref in FP64, or with
// F32 the reference's FP32 arithmetic, (float)(t - tmean) - (float)(tref - tmean)
template <bool F32>
__device__ __forceinline__ double point_dt(double t, double tref, double tmean) {
    if (F32) return (double)((float)(t - tmean) - (float)(tref - tmean));
    return t - tref;
}

// grid cell U (first of the 2m+1 it touches) of a point at time offset dt in
// node set g (x = g df dt); yc = its position in cells (FP32 operations with F32)
template <bool F32>
__device__ __forceinline__ int prep_cell(double dt, int g, const CurveGeom& G, int m,
                                         double& yc) {
    if (F32) {
        const float x = (float)dt * (float)((double)g * G.df);
        float ycf = (x - floorf(x)) * (float)G.n;
        if (ycf >= (float)G.n) ycf -= (float)G.n;   // a fraction just below 1 rounded up
        yc = (double)ycf;
        return (int)floorf(ycf) - m;
    }
    const double x = dt * ((double)g * G.df);
    yc = (x - floor(x)) * (double)G.n;
    return (int)floor(yc) - m;
}

// band-centring modulation exp(2 pi i g Fc dt): the phase reduced in FP64
// before the FP32 sine and cosine, or with F32 formed and reduced in FP32
template <bool F32>
__device__ __forceinline__ float2 prep_phase(double dt, int g, double Fc) {
    if (F32) {
        float r = (float)g * ((float)Fc * (float)dt);
        r -= rintf(r);
        float sn, cs;
        sincospif(2.0f * r, &sn, &cs);
        return make_float2(cs, sn);
    }
    return expi_rev((double)g * (Fc * dt));
}

// weights of a point as FP64: (w, yw) for lomb, (-, y) for nfft
template <int KIND>
__device__ __forceinline__ void prep_weights(const PrepVals& p, double ymean, double ybar,
                                             double s0inv, double& w, double& yw) {
    if (KIND == 0) {
        w = s0inv / (p.wv * p.wv);
        yw = w * (p.yv - ymean - ybar);
    } else if (KIND == 1) {
        w = p.wv;
        yw = p.yv;
    } else {
        w = 0.0;
        yw = p.yv;
    }
}

// writes a point's entries into both node sets (one for nfft) and, for the
// low-frequency sums, its prepared weights in original order; cursor(g,
// bucket) returns the point's slot in the bucket order
template <int KIND, bool F32, typename Cursor>
__device__ __forceinline__ void prep_emit(const PointArrays& pa, const CurveGeom& G, int j,
                                          double dt, double w, double yw, double Fc, int m,
                                          Cursor&& cursor) {
    constexpr int NG = (KIND == 2) ? 1 : 2;
    const bool lowfix = KIND != 2 && pa.dt != nullptr;
    const bool ok = isfinite(dt);
    if (lowfix) {
        pa.w[G.pts + j] = ok ? (float)w : 0.f;
        pa.yw[G.pts + j] = ok ? (float)yw : 0.f;
        pa.dt[G.pts + j] = ok ? dt : 0.0;
    }
    if (!ok) return;   // non-finite times contribute nothing
    const int n2 = 1 << G.p2;
#pragma unroll
    for (int g = 1; g <= NG; ++g) {
        double yc;
        const int U = prep_cell<F32>(dt, g, G, m, yc);
        const float fr = (float)(yc - floor(yc));
        const float2 e = prep_phase<F32>(dt, g, Fc);
        const int pos = cursor(g - 1, (U & (n2 - 1)) / G.c1);
        const int64_t o = G.pts + pos;
        auto rec = [&](double s) {
            return make_int4(U, __float_as_int(fr), __float_as_int((float)s * e.x),
                             __float_as_int((float)s * e.y));
        };
        if (g == 1) {
            pa.nd0[o] = rec(yw);
            if (KIND != 2) pa.nd1[o] = rec(w);
        } else {
            pa.nd2[o] = rec(w);
        }
    }
}

// In-place exclusive scan of h[0 .. nb) by the whole block; returns the total.
template <int NTP>
__device__ int block_exclusive_scan(int* h, int nb, int* wsum) {
    const int tid = threadIdx.x;
    const int per = (nb + NTP - 1) / NTP;
    const int lo = min(nb, tid * per), hi = min(nb, lo + per);
    int local = 0;
    for (int i = lo; i < hi; ++i) local += h[i];
    // inclusive scan of the chunk sums over the block
    const int lane = tid & (kWave - 1), wv = tid / kWave;
    int v = local;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    if (lane == kWave - 1) wsum[wv] = v;
    __syncthreads();
    if (tid == 0) {
        int acc = 0;
        for (int w = 0; w < NTP / kWave; ++w) {
            const int x = wsum[w];
            wsum[w] = acc;
            acc += x;
        }
        wsum[NTP / kWave] = acc;
    }
    __syncthreads();
    int run = v - local + wsum[wv];   // exclusive prefix of this chunk
    for (int i = lo; i < hi; ++i) {
        const int x = h[i];
        h[i] = run;
        run += x;
    }
    const int total = wsum[NTP / kWave];
    __syncthreads();
    return total;
}

// NTP threads per curve; curves taken from list[blockIdx.x] (256 threads for
// most curves, 1024 for long ones without the chunked path)
template <int KIND, int NTP, bool F32>
__global__ __launch_bounds__(NTP) void prep_kernel(const CurveGeom* __restrict__ geom,
                                                   const int32_t* __restrict__ list,
                                                   CurveState* __restrict__ state,
                                                   const double* __restrict__ t,
                                                   const void* __restrict__ a,
                                                   const void* __restrict__ b,
                                                   PointArrays pa, int32_t* __restrict__ bkt,
                                                   int m, int method, int y_is_f32) {
    __shared__ double red[NTP / kWave];
    __shared__ int hist[2][4096];
    __shared__ int wsum[NTP / kWave + 1];
    const int c = list[blockIdx.x];
    const CurveGeom G = geom[c];
    const int tid = threadIdx.x;
    const int N = G.n_obs;

    // ---- phase A: reductions ------------------------------------------------
    double tmin = INFINITY, tmax = -INFINITY, sy = 0.0, cnt = 0.0, s0 = 0.0, st = 0.0, ct = 0.0;
#pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
        PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
        tmin = fmin(tmin, p.t);
        tmax = fmax(tmax, p.t);
        if (F32 && isfinite(p.t)) {
            st += p.t;
            ct += 1.0;
        }
        if (KIND == 0) {
            if (isfinite(p.yv)) {
                sy += p.yv;
                cnt += 1.0;
            }
            s0 += 1.0 / (p.wv * p.wv);
        }
    }
    tmin = block_min<NTP>(tmin, red);
    tmax = block_max<NTP>(tmax, red);
    double tmean = 0.0;
    if (F32) {
        st = block_sum<NTP>(st, red);
        ct = block_sum<NTP>(ct, red);
        tmean = st / ct;
    }
    double ymean = 0.0, ybar = G.ybar_in, yy = G.yy_in, s0inv = 1.0;
    if (KIND == 0) {
        sy = block_sum<NTP>(sy, red);
        cnt = block_sum<NTP>(cnt, red);
        s0 = block_sum<NTP>(s0, red);
        ymean = sy / cnt;
        s0inv = 1.0 / s0;
        double s1 = 0.0;
    #pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            s1 += (p.yv - ymean) / (p.wv * p.wv);
        }
        ybar = block_sum<NTP>(s1, red) * s0inv;
        double s2 = 0.0;
    #pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            const double d = p.yv - ymean - ybar;
            s2 += d * d / (p.wv * p.wv);
        }
        yy = block_sum<NTP>(s2, red) * s0inv;
    }
    if (tid == 0) {
        CurveState st;
        st.tref = tmin;
        st.ybar = ybar;
        st.yy = yy;
        st.tspan = tmax - tmin;
        st.tmean = tmean;
        state[c] = st;
    }
    const double tref = tmin;

    auto weights_of = [&](const PrepVals& p, double& w, double& yw) {
        prep_weights<KIND>(p, ymean, ybar, s0inv, w, yw);
    };

    if (method == GLS_METHOD_DIRECT || method == GLS_METHOD_DIRECT_CUSTOM) {
    #pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            double w, yw;
            weights_of(p, w, yw);
            pa.w[G.pts + j] = (float)w;
            pa.yw[G.pts + j] = (float)yw;
            pa.dt[G.pts + j] = p.t - tref;
        }
        return;
    }

    // ---- phase B: cells, modulation and bucketing of both node sets --------
    // (x = df (t - tref) for SYW and SW(f), 2x for SW(2f)); one counting pass,
    // one block scan per set, one scatter pass
    const int p2 = G.p2;
    const int n2 = 1 << p2;
    const int c1 = G.c1;
    const int nb = n2 / c1;
    const double Fc = G.f0 + (double)G.kc * G.df;
    constexpr int NG = (KIND == 2) ? 1 : 2;
    for (int i = tid; i < nb; i += NTP) {
        hist[0][i] = 0;
        hist[1][i] = 0;
    }
    __syncthreads();
    auto cell_of = [&](double dt, int g, double& yc) -> int {
        return prep_cell<F32>(dt, g, G, m, yc);
    };
#pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
        const double dt = point_dt<F32>(t[G.obs + j], tref, tmean);
        if (!isfinite(dt)) continue;  // non-finite times contribute nothing
#pragma unroll
        for (int g = 1; g <= NG; ++g) {
            double yc;
            const int U = cell_of(dt, g, yc);
            atomicAdd(&hist[g - 1][(U & (n2 - 1)) / c1], 1);
        }
    }
    __syncthreads();
#pragma unroll
    for (int g = 1; g <= NG; ++g) {
        int32_t* bk = bkt + G.bkt + (int64_t)(g - 1) * (nb + 1);
        const int total = block_exclusive_scan<NTP>(hist[g - 1], nb, wsum);
        for (int i = tid; i < nb; i += NTP) bk[i] = hist[g - 1][i];
        if (tid == 0) bk[nb] = total;
    }
    __syncthreads();
#pragma unroll 4
    for (int j = tid; j < N; j += NTP) {
        PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
        double w, yw;
        weights_of(p, w, yw);
        prep_emit<KIND, F32>(pa, G, j, point_dt<F32>(p.t, tref, tmean), w, yw, Fc, m,
                             [&](int gs, int bucket) { return atomicAdd(&hist[gs][bucket], 1); });
    }
}

// Wide curves (more than kWidePrep points, NFFT method) are prepared by
// kernels over chunks of kWideChunk points: blockIdx.y picks the curve
// (wide[blockIdx.y], record aux[blockIdx.y], chunk partials
// part[blockIdx.y * gridDim.x ..]), blockIdx.x its chunk.  The grid has the
// same chunk count for every wide curve (that of the longest); chunks past a
// curve's end are empty.
//
// wide_clear_kernel (one block per wide curve) zeroes the record's counters
// and accumulators at every execute.
__global__ __launch_bounds__(NT) void wide_clear_kernel(const CurveGeom* __restrict__ geom,
                                                        const int32_t* __restrict__ wide,
                                                        WideAux* __restrict__ aux) {
    const CurveGeom& G = geom[wide[blockIdx.x]];
    WideAux& X = aux[blockIdx.x];
    const int tid = threadIdx.x;
    const int nb = (1 << G.p2) / G.c1;
    if (tid == 0) {
        X.done = 0;
        X.red_done = 0;
        X.yy_done = 0;
    }
    for (int i = tid; i < nb; i += NT) {
        X.hist[0][i] = 0;
        X.hist[1][i] = 0;
    }
    for (int i = tid; i < kLowMax; i += NT) X.low_done[i] = 0;
}

__device__ __forceinline__ double agent_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// wide_reduce_kernel: the reductions of phase A per chunk (time range; for
// raw input the sums of the finite y, of 1/dy^2 and y/dy^2; with FP32
// coordinates the sum of the finite t), each chunk's partials into its slot;
// the last block to finish combines the slots in chunk order and writes the
// curve's state (tref, span, ybar; yy follows from the scatter) and the
// weight constants of the record.
template <int KIND, bool F32>
__global__ __launch_bounds__(NT) void wide_reduce_kernel(const CurveGeom* __restrict__ geom,
                                                         const int32_t* __restrict__ wide,
                                                         CurveState* __restrict__ state,
                                                         const double* __restrict__ t,
                                                         const void* __restrict__ a,
                                                         const void* __restrict__ b,
                                                         WideAux* __restrict__ aux,
                                                         WidePart* __restrict__ part,
                                                         int y_is_f32) {
    __shared__ double red[NT / kWave];
    __shared__ int last;
    const int c = wide[blockIdx.y];
    const CurveGeom G = geom[c];
    WideAux& X = aux[blockIdx.y];
    WidePart* P = part + (int64_t)blockIdx.y * gridDim.x;
    const int tid = threadIdx.x;
    WidePart v{};
    v.tmin = INFINITY;
    v.tmax = -INFINITY;
    const int j1 = min(G.n_obs, (int)(blockIdx.x + 1) * kWideChunk);
    for (int j = (int)blockIdx.x * kWideChunk + tid; j < j1; j += NT) {
        PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
        v.tmin = fmin(v.tmin, p.t);
        v.tmax = fmax(v.tmax, p.t);
        if (F32 && isfinite(p.t)) {
            v.st += p.t;
            v.ct += 1.0;
        }
        if (KIND == 0) {
            if (isfinite(p.yv)) {
                v.sy += p.yv;
                v.cy += 1.0;
            }
            const double iw = 1.0 / (p.wv * p.wv);
            v.sw += iw;
            v.swy += p.yv * iw;
        }
    }
    auto reduce = [&](WidePart& q) {
        q.tmin = block_min<NT>(q.tmin, red);
        q.tmax = block_max<NT>(q.tmax, red);
        if (F32) {
            q.st = block_sum<NT>(q.st, red);
            q.ct = block_sum<NT>(q.ct, red);
        }
        if (KIND == 0) {
            q.sy = block_sum<NT>(q.sy, red);
            q.cy = block_sum<NT>(q.cy, red);
            q.sw = block_sum<NT>(q.sw, red);
            q.swy = block_sum<NT>(q.swy, red);
        }
    };
    reduce(v);
    if (tid == 0) P[blockIdx.x] = v;
    __threadfence();
    __syncthreads();
    if (tid == 0) last = (atomicAdd(&X.red_done, 1) == (int)gridDim.x - 1);
    __syncthreads();
    if (!last) return;   // block-uniform
    __threadfence();
    WidePart q{};
    q.tmin = INFINITY;
    q.tmax = -INFINITY;
    for (int i = tid; i < (int)gridDim.x; i += NT) {
        const double* s = &P[i].tmin;
        q.tmin = fmin(q.tmin, agent_load(s + 0));
        q.tmax = fmax(q.tmax, agent_load(s + 1));
        q.st += agent_load(s + 2);
        q.ct += agent_load(s + 3);
        q.sy += agent_load(s + 4);
        q.cy += agent_load(s + 5);
        q.sw += agent_load(s + 6);
        q.swy += agent_load(s + 7);
    }
    reduce(q);
    if (tid == 0) {
        CurveState st;
        st.tref = q.tmin;
        st.tspan = q.tmax - q.tmin;
        st.tmean = F32 ? q.st / q.ct : 0.0;
        if (KIND == 0) {
            const double ymean = q.sy / q.cy;
            X.ymean = ymean;
            X.s0inv = 1.0 / q.sw;
            st.ybar = q.swy / q.sw - ymean;
            st.yy = 0.0;   // the scatter's last block writes it
        } else {
            X.ymean = 0.0;
            X.s0inv = 1.0;
            st.ybar = G.ybar_in;
            st.yy = G.yy_in;
        }
        state[c] = st;
    }
}

// Bucketing of the wide curves' points over many blocks: blockIdx.y picks the
// curve (wide[blockIdx.y], record aux[blockIdx.y]), blockIdx.x its chunk of
// kWideChunk points.  Each block counts its chunk in LDS and adds the
// non-zero counts to the record with one global atomic per bucket (per-point
// global atomics on a few hundred counters serialise on their cache lines).
// Counting: the last block of the curve to finish (the block count is
// gridDim.x for every curve) scans the totals into the bucket starts bkt and
// the cursors.  Scatter: each block reserves its slots per bucket with one
// atomic on the cursor, then ranks its points in LDS.
template <int KIND, bool F32>
__device__ __forceinline__ void wide_chunk_counts(int (&h)[2][kMaxBuckets], const CurveGeom& G,
                                                  const double* __restrict__ t, double tref,
                                                  double tmean, int m) {
    constexpr int NG = (KIND == 2) ? 1 : 2;
    const int tid = threadIdx.x;
    const int n2 = 1 << G.p2;
    const int nb = n2 / G.c1;
    for (int i = tid; i < nb; i += NT) {
        h[0][i] = 0;
        h[1][i] = 0;
    }
    __syncthreads();
    const int j1 = min(G.n_obs, (int)(blockIdx.x + 1) * kWideChunk);
    for (int j = (int)blockIdx.x * kWideChunk + tid; j < j1; j += NT) {
        const double dt = point_dt<F32>(t[G.obs + j], tref, tmean);
        if (!isfinite(dt)) continue;
#pragma unroll
        for (int g = 1; g <= NG; ++g) {
            double yc;
            const int U = prep_cell<F32>(dt, g, G, m, yc);
            atomicAdd(&h[g - 1][(U & (n2 - 1)) / G.c1], 1);
        }
    }
    __syncthreads();
}

template <int KIND, bool F32>
__global__ __launch_bounds__(NT) void wide_count_kernel(const CurveGeom* __restrict__ geom,
                                                        const int32_t* __restrict__ wide,
                                                        const CurveState* __restrict__ state,
                                                        const double* __restrict__ t,
                                                        WideAux* __restrict__ aux,
                                                        int32_t* __restrict__ bkt, int m) {
    __shared__ int h[2][kMaxBuckets];
    __shared__ int wsum[NT / kWave + 1];
    __shared__ int last;
    constexpr int NG = (KIND == 2) ? 1 : 2;
    const int tid = threadIdx.x;
    const CurveGeom G = geom[wide[blockIdx.y]];
    WideAux& X = aux[blockIdx.y];
    const int nb = (1 << G.p2) / G.c1;
    const CurveState st = state[wide[blockIdx.y]];
    wide_chunk_counts<KIND, F32>(h, G, t, st.tref, st.tmean, m);
#pragma unroll
    for (int g = 0; g < NG; ++g)
        for (int i = tid; i < nb; i += NT)
            if (h[g][i] != 0) atomicAdd(&X.hist[g][i], h[g][i]);
    __threadfence();
    __syncthreads();
    if (tid == 0) last = (atomicAdd(&X.done, 1) == (int)gridDim.x - 1);
    __syncthreads();
    if (!last) return;   // block-uniform
    __threadfence();
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        for (int i = tid; i < nb; i += NT)
            h[g][i] = __hip_atomic_load(&X.hist[g][i], __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const int total = block_exclusive_scan<NT>(h[g], nb, wsum);
        int32_t* bk = bkt + G.bkt + (int64_t)g * (nb + 1);
        for (int i = tid; i < nb; i += NT) {
            bk[i] = h[g][i];
            X.hist[g][i] = h[g][i];
        }
        if (tid == 0) bk[nb] = total;
        __syncthreads();
    }
}

// The scatter also sums (y - ymean - ybar)^2 / dy^2 per chunk (raw input);
// its last block combines the chunks in order into the curve's yy.
template <int KIND, bool F32>
__global__ __launch_bounds__(NT) void wide_scatter_kernel(const CurveGeom* __restrict__ geom,
                                                          const int32_t* __restrict__ wide,
                                                          CurveState* __restrict__ state,
                                                          const double* __restrict__ t,
                                                          const void* __restrict__ a,
                                                          const void* __restrict__ b,
                                                          PointArrays pa,
                                                          WideAux* __restrict__ aux,
                                                          WidePart* __restrict__ part, int m,
                                                          int y_is_f32) {
    __shared__ int h[2][kMaxBuckets];
    __shared__ double red[NT / kWave];
    __shared__ int last;
    constexpr int NG = (KIND == 2) ? 1 : 2;
    const int tid = threadIdx.x;
    const CurveGeom G = geom[wide[blockIdx.y]];
    WideAux& X = aux[blockIdx.y];
    const CurveState st = state[wide[blockIdx.y]];
    const int nb = (1 << G.p2) / G.c1;
    wide_chunk_counts<KIND, F32>(h, G, t, st.tref, st.tmean, m);
    // this block's first slot in each bucket
#pragma unroll
    for (int g = 0; g < NG; ++g)
        for (int i = tid; i < nb; i += NT)
            if (h[g][i] != 0) h[g][i] = atomicAdd(&X.hist[g][i], h[g][i]);
    __syncthreads();
    const double ymean = X.ymean, s0inv = X.s0inv;
    const double Fc = G.f0 + (double)G.kc * G.df;
    const int j1 = min(G.n_obs, (int)(blockIdx.x + 1) * kWideChunk);
    double syy = 0.0;
    for (int j = (int)blockIdx.x * kWideChunk + tid; j < j1; j += NT) {
        PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
        double w, yw;
        prep_weights<KIND>(p, ymean, st.ybar, s0inv, w, yw);
        if (KIND == 0) {
            const double d = p.yv - ymean - st.ybar;
            syy += d * d / (p.wv * p.wv);
        }
        prep_emit<KIND, F32>(pa, G, j, point_dt<F32>(p.t, st.tref, st.tmean), w, yw, Fc, m,
                             [&](int gs, int bucket) { return atomicAdd(&h[gs][bucket], 1); });
    }
    if (KIND != 0) return;
    WidePart* P = part + (int64_t)blockIdx.y * gridDim.x;
    syy = block_sum<NT>(syy, red);
    if (tid == 0) P[blockIdx.x].syy = syy;
    __threadfence();
    __syncthreads();
    if (tid == 0) last = (atomicAdd(&X.yy_done, 1) == (int)gridDim.x - 1);
    __syncthreads();
    if (!last) return;   // block-uniform
    __threadfence();
    double acc = 0.0;
    for (int i = tid; i < (int)gridDim.x; i += NT) acc += agent_load(&P[i].syy);
    acc = block_sum<NT>(acc, red);
    if (tid == 0) state[wide[blockIdx.y]].yy = acc * s0inv;
}

}  // namespace

namespace {
template <bool F32>
hipError_t launch_prep_t(const CurveGeom* d_geom, const int32_t* small, int nsmall,
                         const int32_t* wide, int nwide, CurveState* state, const double* t,
                         const void* a, const void* b, PointArrays pa, int32_t* bkt,
                         const PlanParams& pp, int y_is_f32, WideAux* aux, WidePart* part,
                         int wide_max_obs, hipStream_t s) {
    const int kind = (pp.kind == 1) ? 2 : (pp.input_kind == GLS_INPUT_RAW ? 0 : 1);
    const int method = (pp.kind == 1) ? GLS_METHOD_NFFT : pp.method;
    const int yf = (pp.kind == 1) ? y_is_f32 : 0;
#define GLS_PREP_LAUNCH(K, NTP, LIST, COUNT)                                                  \
    hipLaunchKernelGGL((prep_kernel<K, NTP, F32>), dim3(COUNT), dim3(NTP), 0, s, d_geom,      \
                       LIST, state, t, a, b, pa, bkt, pp.m, method, yf)
    if (nsmall > 0) {
        if (kind == 0) GLS_PREP_LAUNCH(0, 256, small, nsmall);
        else if (kind == 1) GLS_PREP_LAUNCH(1, 256, small, nsmall);
        else GLS_PREP_LAUNCH(2, 256, small, nsmall);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (nwide <= 0) return hipGetLastError();
    const bool split = aux != nullptr && part != nullptr && method == GLS_METHOD_NFFT &&
                       wide_max_obs > 0;
    if (!split) {
        if (kind == 0) GLS_PREP_LAUNCH(0, 1024, wide, nwide);
        else if (kind == 1) GLS_PREP_LAUNCH(1, 1024, wide, nwide);
        else GLS_PREP_LAUNCH(2, 1024, wide, nwide);
        return hipGetLastError();
    }
#undef GLS_PREP_LAUNCH
    hipLaunchKernelGGL(wide_clear_kernel, dim3(nwide), dim3(NT), 0, s, d_geom, wide, aux);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)((wide_max_obs + kWideChunk - 1) / kWideChunk), (unsigned)nwide);
#define GLS_WIDE_LAUNCH(K)                                                                    \
    hipLaunchKernelGGL((wide_reduce_kernel<K, F32>), grid, dim3(NT), 0, s, d_geom, wide,       \
                       state, t, a, b, aux, part, yf);                                        \
    hipLaunchKernelGGL((wide_count_kernel<K, F32>), grid, dim3(NT), 0, s, d_geom, wide, state, \
                       t, aux, bkt, pp.m);                                                    \
    hipLaunchKernelGGL((wide_scatter_kernel<K, F32>), grid, dim3(NT), 0, s, d_geom, wide,      \
                       state, t, a, b, pa, aux, part, pp.m, yf)
    if (kind == 0) {
        GLS_WIDE_LAUNCH(0);
    } else if (kind == 1) {
        GLS_WIDE_LAUNCH(1);
    } else {
        GLS_WIDE_LAUNCH(2);
    }
#undef GLS_WIDE_LAUNCH
    return hipGetLastError();
}
}  // namespace

hipError_t launch_prep(const CurveGeom* d_geom, const int32_t* small, int nsmall,
                       const int32_t* wide, int nwide, CurveState* state, const double* t,
                       const void* a, const void* b, PointArrays pa, int32_t* bkt,
                       const PlanParams& pp, int y_is_f32, WideAux* aux, WidePart* part,
                       int wide_max_obs, hipStream_t s) {
    if (pp.fp32_coords && pp.kind == 0)
        return launch_prep_t<true>(d_geom, small, nsmall, wide, nwide, state, t, a, b, pa, bkt,
                                   pp, y_is_f32, aux, part, wide_max_obs, s);
    return launch_prep_t<false>(d_geom, small, nsmall, wide, nwide, state, t, a, b, pa, bkt, pp,
                                y_is_f32, aux, part, wide_max_obs, s);
}

}  // namespace gls
