// Per-curve data preparation and point bucketing.
//
// One workgroup per curve:
//   phase A  reductions: tmin/tmax, and for raw input the weighted mean and
//            variance of LombScargleMemory.setdata (cuvarbase/lombscargle.py:
//            243-257) after the nanmean centring of normalize_light_curves
//            (cuvarbase/utils.py:134-141);
//   phase B  (NFFT method) for each of the two node sets -- x = df (t - tref)
//            for SYW and SW(f), 2x for SW(2f) -- the grid cell U, the in-cell
//            offset and the band-centring modulation exp(2 pi i Fc (t - tref))
//            of every point, in FP64 before the cast to FP32, then a counting
//            sort of the points by pass-1 column block (so every pass-1 tile
//            reads only the points that touch it);
//            (direct method) the FP32 weights and FP64 time offsets.
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

template <int KIND>
__global__ __launch_bounds__(NT) void prep_kernel(const CurveGeom* __restrict__ geom,
                                                  CurveState* __restrict__ state,
                                                  const double* __restrict__ t,
                                                  const void* __restrict__ a,
                                                  const void* __restrict__ b,
                                                  PointArrays pa, int32_t* __restrict__ bkt,
                                                  int m, int method, int y_is_f32) {
    __shared__ double red[NT / kWave];
    __shared__ int hist[4096];
    const int c = blockIdx.x;
    const CurveGeom G = geom[c];
    const int tid = threadIdx.x;
    const int N = G.n_obs;

    // ---- phase A: reductions ------------------------------------------------
    double tmin = INFINITY, tmax = -INFINITY, sy = 0.0, cnt = 0.0, s0 = 0.0;
    for (int j = tid; j < N; j += NT) {
        PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
        tmin = fmin(tmin, p.t);
        tmax = fmax(tmax, p.t);
        if (KIND == 0) {
            if (isfinite(p.yv)) {
                sy += p.yv;
                cnt += 1.0;
            }
            s0 += 1.0 / (p.wv * p.wv);
        }
    }
    tmin = block_min<NT>(tmin, red);
    tmax = block_max<NT>(tmax, red);
    double ymean = 0.0, ybar = G.ybar_in, yy = G.yy_in, s0inv = 1.0;
    if (KIND == 0) {
        sy = block_sum<NT>(sy, red);
        cnt = block_sum<NT>(cnt, red);
        s0 = block_sum<NT>(s0, red);
        ymean = sy / cnt;
        s0inv = 1.0 / s0;
        double s1 = 0.0;
        for (int j = tid; j < N; j += NT) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            s1 += (p.yv - ymean) / (p.wv * p.wv);
        }
        ybar = block_sum<NT>(s1, red) * s0inv;
        double s2 = 0.0;
        for (int j = tid; j < N; j += NT) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            const double d = p.yv - ymean - ybar;
            s2 += d * d / (p.wv * p.wv);
        }
        yy = block_sum<NT>(s2, red) * s0inv;
    }
    if (tid == 0) {
        CurveState st;
        st.tref = tmin;
        st.ybar = ybar;
        st.yy = yy;
        st.tspan = tmax - tmin;
        state[c] = st;
    }
    const double tref = tmin;

    // weights of point j as FP64: (w, yw) for lomb, (y, -) for nfft
    auto weights_of = [&](const PrepVals& p, double& w, double& yw) {
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
    };

    if (method == GLS_METHOD_DIRECT || method == GLS_METHOD_DIRECT_CUSTOM) {
        for (int j = tid; j < N; j += NT) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            double w, yw;
            weights_of(p, w, yw);
            pa.w[G.pts + j] = (float)w;
            pa.yw[G.pts + j] = (float)yw;
            pa.dt[G.pts + j] = p.t - tref;
        }
        return;
    }

    // prepared points for the low-frequency direct sums of lomb NFFT plans
    if (KIND != 2 && pa.dt != nullptr) {
        for (int j = tid; j < N; j += NT) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            double w, yw;
            weights_of(p, w, yw);
            const double dt = p.t - tref;
            pa.w[G.pts + j] = isfinite(dt) ? (float)w : 0.f;
            pa.yw[G.pts + j] = isfinite(dt) ? (float)yw : 0.f;
            pa.dt[G.pts + j] = isfinite(dt) ? dt : 0.0;
        }
    }

    // ---- phase B: cells, modulation and bucketing -------------------------
    const int n = G.n;
    const int p2 = G.p2;
    const int n2 = 1 << p2;
    const int c1 = G.c1;
    const int nb = n2 / c1;
    const double Fc = G.f0 + (double)G.kc * G.df;
    const int ngtypes = (KIND == 2) ? 1 : 2;
    for (int g = 1; g <= ngtypes; ++g) {
        int32_t* bk = bkt + G.bkt + (int64_t)(g - 1) * (nb + 1);
        for (int i = tid; i < nb; i += NT) hist[i] = 0;
        __syncthreads();
        const double gdf = (double)g * G.df;
        for (int j = tid; j < N; j += NT) {
            const double dt = t[G.obs + j] - tref;
            if (!isfinite(dt)) continue;  // non-finite times contribute nothing
            const double x = dt * gdf;
            const double y = (x - floor(x)) * (double)n;
            const int U = (int)floor(y) - m;
            atomicAdd(&hist[(U & (n2 - 1)) / c1], 1);
        }
        __syncthreads();
        if (tid == 0) {
            // exclusive scan (nb <= 4096, a few microseconds at most)
            int acc = 0;
            for (int i = 0; i < nb; ++i) {
                const int h = hist[i];
                hist[i] = acc;
                bk[i] = acc;
                acc += h;
            }
            bk[nb] = acc;
        }
        __syncthreads();
        for (int j = tid; j < N; j += NT) {
            PrepVals p = load_point<KIND>(t, a, b, G.obs + j, y_is_f32);
            const double dt = p.t - tref;
            if (!isfinite(dt)) continue;
            const double x = dt * gdf;
            const double yc = (x - floor(x)) * (double)n;
            const double fl = floor(yc);
            const int U = (int)fl - m;
            const float fr = (float)(yc - fl);
            const float2 e = expi_rev((double)g * (Fc * dt));
            double w, yw;
            weights_of(p, w, yw);
            const int pos = atomicAdd(&hist[(U & (n2 - 1)) / c1], 1);
            const int64_t o = G.pts + pos;
            if (g == 1) {
                pa.U1[o] = U;
                pa.fr1[o] = fr;
                pa.v0[o] = make_float2((float)yw * e.x, (float)yw * e.y);
                if (KIND != 2) pa.v1[o] = make_float2((float)w * e.x, (float)w * e.y);
            } else {
                pa.U2[o] = U;
                pa.fr2[o] = fr;
                pa.v2[o] = make_float2((float)w * e.x, (float)w * e.y);
            }
        }
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_prep(const CurveGeom* d_geom, int ncurves, CurveState* state,
                       const double* t, const void* a, const void* b, PointArrays pa,
                       int32_t* bkt, const PlanParams& pp, int y_is_f32, hipStream_t s) {
    if (ncurves <= 0) return hipSuccess;
    dim3 grid(ncurves), block(NT);
    if (pp.kind == 1) {
        hipLaunchKernelGGL(prep_kernel<2>, grid, block, 0, s, d_geom, state, t, a, b, pa, bkt,
                           pp.m, GLS_METHOD_NFFT, y_is_f32);
    } else if (pp.input_kind == GLS_INPUT_RAW) {
        hipLaunchKernelGGL(prep_kernel<0>, grid, block, 0, s, d_geom, state, t, a, b, pa, bkt,
                           pp.m, pp.method, 0);
    } else {
        hipLaunchKernelGGL(prep_kernel<1>, grid, block, 0, s, d_geom, state, t, a, b, pa, bkt,
                           pp.m, pp.method, 0);
    }
    return hipGetLastError();
}

}  // namespace gls
