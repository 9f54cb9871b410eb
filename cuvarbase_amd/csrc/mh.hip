// Multi-harmonic generalized Lomb-Scargle power from harmonic sums.
//
// For H harmonics the floating-mean model y ~ c0 + sum_h (a_h cos h phi + b_h sin h phi)
// needs, at each frequency f, the weighted sums
//   c[n] + i s[n] = sum_j w_j exp(i n phi_j),  n = 1 .. 2H,   phi_j = 2 pi f t_j
//   yc[n] + i ys[n] = sum_j yw_j exp(i n phi_j),  n = 1 .. H,
// i.e. 3H adjoint NFFTs of one curve, which the NFFT plan computes.  This
// kernel forms, per frequency, the matrices of mhdirect_sums
// (cuvarbase/lombscargle.py:337-400) from the product-to-sum identities,
//   CC[n][m] = (c[n+m] + c[|n-m|])/2 - C_n C_m
//   CS[n][m] = (s[n+m] - sgn(n-m) s[|n-m|])/2 - C_n S_m
//   SS[n][m] = (c[|n-m|] - c[n+m])/2 - S_n S_m,
// adds the amplitude prior to the diagonals (add_regularization, :403-447),
// solves A theta = b by Cholesky in FP64 and returns P = theta . b / YY, which
// equals the reference's (2 theta.b - theta^T A theta) / YY at the solution
// (mhgls_from_sums, :497-536).  Invalid results are the -1 sentinel.
#include "device_util.hpp"

namespace gls {

namespace {

constexpr int NTM = 256;

template <int H, int OUT>
__global__ __launch_bounds__(NTM) void mh_power_kernel(const float2* __restrict__ sums,
                                                       const gls_mh_curve* __restrict__ curves,
                                                       void* __restrict__ out) {
    constexpr int D = 2 * H;
    const gls_mh_curve c = curves[blockIdx.y];
    const int k = blockIdx.x * NTM + threadIdx.x;
    if (k >= c.nf) return;
    const int64_t K = c.nf;
    const float2* S = sums + c.sums_offset;

    double cw[D + 1], sw[D + 1];   // c[n], s[n], n = 0 .. 2H
    cw[0] = c.sum_w;
    sw[0] = 0.0;
#pragma unroll
    for (int n = 1; n <= D; ++n) {
        const float2 z = S[(int64_t)(H + n - 1) * K + k];
        cw[n] = (double)z.x;
        sw[n] = (double)z.y;
    }
    double b[D];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
        const float2 z = S[(int64_t)(n - 1) * K + k];
        b[n - 1] = (double)z.x - c.ybar * cw[n];
        b[H + n - 1] = (double)z.y - c.ybar * sw[n];
    }
    // A = [[CC + lam, CS], [CS^T, SS + lam]]
    double A[D][D];
#pragma unroll
    for (int n = 1; n <= H; ++n) {
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const int dp = n + m;
            const int dm = n > m ? n - m : m - n;
            const double sg = (n - m) >= 0 ? 1.0 : -1.0;
            const double cc = 0.5 * (cw[dp] + cw[dm]) - cw[n] * cw[m];
            const double cs = 0.5 * (sw[dp] - sg * sw[dm]) - cw[n] * sw[m];
            const double ss = 0.5 * (cw[dm] - cw[dp]) - sw[n] * sw[m];
            A[n - 1][m - 1] = cc + (n == m ? c.lam : 0.0);
            A[H + n - 1][H + m - 1] = ss + (n == m ? c.lam : 0.0);
            A[n - 1][H + m - 1] = cs;
            A[H + m - 1][n - 1] = cs;
        }
    }
    // Cholesky A = L L^T (in place, lower triangle), then L y = b, L^T theta = y
    bool ok = true;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        double d = A[j][j];
#pragma unroll
        for (int q = 0; q < j; ++q) d -= A[j][q] * A[j][q];
        ok = ok && d > 0.0;
        const double ljj = sqrt(d > 0.0 ? d : 1.0);
        A[j][j] = ljj;
#pragma unroll
        for (int i = j + 1; i < D; ++i) {
            double v = A[i][j];
#pragma unroll
            for (int q = 0; q < j; ++q) v -= A[i][q] * A[j][q];
            A[i][j] = v / ljj;
        }
    }
    double y[D];
#pragma unroll
    for (int i = 0; i < D; ++i) {
        double v = b[i];
#pragma unroll
        for (int q = 0; q < i; ++q) v -= A[i][q] * y[q];
        y[i] = v / A[i][i];
    }
    // theta . b = y . y  (b = L y, theta = L^-T y)
    double tb = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) tb += y[i] * y[i];
    double P = tb / c.yy;
    if (!ok || !isfinite(P) || P < 0.0) P = -1.0;
    if (OUT == 0) ((float*)out)[c.out_offset + k] = (float)P;
    else ((double*)out)[c.out_offset + k] = P;
}

template <int H>
hipError_t mh_dispatch(const float2* sums, const gls_mh_curve* curves, int ncurves, int max_nf,
                       void* power, int out_f64, hipStream_t s) {
    const dim3 grid((max_nf + NTM - 1) / NTM, ncurves);
    if (out_f64)
        hipLaunchKernelGGL((mh_power_kernel<H, 1>), grid, dim3(NTM), 0, s, sums, curves, power);
    else
        hipLaunchKernelGGL((mh_power_kernel<H, 0>), grid, dim3(NTM), 0, s, sums, curves, power);
    return hipGetLastError();
}

}  // namespace

}  // namespace gls

using namespace gls;

extern "C" int gls_mh_power(const void* sums, const gls_mh_curve* curves, int32_t ncurves,
                            int32_t nharm, int32_t max_nf, void* power, int32_t power_is_f64,
                            void* stream) {
    int rc = require_init();
    if (rc) return rc;
    if (nharm < 1 || nharm > 4) return fail(GLS_E_UNSUPPORTED, "nharm must be in [1, 4]");
    if (ncurves < 0 || ncurves > 65535) return fail(GLS_E_INVALID, "ncurves must be in [0, 65535]");
    if (ncurves == 0 || max_nf <= 0) return GLS_OK;
    if (!sums || !curves || !power) return fail(GLS_E_INVALID, "NULL device pointer");
    const hipStream_t s = (hipStream_t)stream;
    const float2* z = (const float2*)sums;
    const int f64 = power_is_f64 ? 1 : 0;
    switch (nharm) {
        case 1: GLS_HIP_TRY(mh_dispatch<1>(z, curves, ncurves, max_nf, power, f64, s)); break;
        case 2: GLS_HIP_TRY(mh_dispatch<2>(z, curves, ncurves, max_nf, power, f64, s)); break;
        case 3: GLS_HIP_TRY(mh_dispatch<3>(z, curves, ncurves, max_nf, power, f64, s)); break;
        default: GLS_HIP_TRY(mh_dispatch<4>(z, curves, ncurves, max_nf, power, f64, s)); break;
    }
    return GLS_OK;
}
