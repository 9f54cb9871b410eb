// Grid-size catalogue of the four-step NFFT grids: n = n1 * n2 with n2 = 2^p2
// (pass-2 row length, 64..4096) and n1 (pass-1 column length) from the list
// below.  The list was chosen with the cost model n * (1 + 0.12 (S(n1) +
// S(n2))), S = number of LDS radix stages, over grid sizes 2^10 .. 2^22: mean
// overshoot of n over the requested sigma*nf is ~7% (powers of two: ~44%).
#pragma once

#define GLS_PASS1_SIZES(X)                                                         \
    X(16) X(20) X(24) X(25) X(27) X(36) X(40) X(48) X(72) X(80) X(81) X(128)      \
    X(144) X(200) X(216) X(256) X(384) X(400) X(405) X(432) X(640) X(648) X(768)  \
    X(1080) X(1152) X(1280) X(1296) X(2048) X(2160) X(2304) X(3072) X(3200)        \
    X(3240) X(3456) X(4096)

namespace gls {
#define GLS_SIZE_ENTRY(v) v,
constexpr int kPass1Sizes[] = {GLS_PASS1_SIZES(GLS_SIZE_ENTRY)};
#undef GLS_SIZE_ENTRY
constexpr int kNumPass1Sizes = sizeof(kPass1Sizes) / sizeof(int);
constexpr int kMinLog2Row = 6;    // n2 >= 64
constexpr int kMaxLog2Row = 12;   // n2 <= 4096
}  // namespace gls
