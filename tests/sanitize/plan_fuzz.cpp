// Randomised descriptions through dxg2_plan_dry_run, the host half of dxg2_plan_create,
// built with AddressSanitizer and UndefinedBehaviorSanitizer: every work-list builder
// runs over shapes, frame counts, dtypes, flags and bin patterns (empty bins, one-pixel
// bins, thousands of bins, excluded pixels, sparse bin numbers), and the summary it
// reports is checked against counts made here.  Exit status 0 = all cases passed.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "dxg2.h"

namespace {

constexpr int64_t kBlock = 64;      // pixels per k-block
constexpr int64_t kRowAlign = 256;  // slab rows are nframes rounded up to this

int fail(int c, const char* what, int64_t got, int64_t want) {
    std::fprintf(stderr, "case %d: %s = %lld, expected %lld\n", c, what, (long long)got,
                 (long long)want);
    return 1;
}

int check_case(int c, std::mt19937_64& rng) {
    std::uniform_int_distribution<int> U(0, 1 << 30);
    const int H = 1 + U(rng) % 160, W = 1 + U(rng) % 160;
    const int shapes[] = {1, 2, 63, 64, 65, 255, 256, 257, 1000, 4097, 20000, 100000};
    const int N = (U(rng) % 3 == 0) ? 1 + U(rng) % 3000 : shapes[U(rng) % 12];
    const int dtypes[] = {DXG2_U8, DXG2_I8, DXG2_U16};
    const int dtype = dtypes[U(rng) % 3];
    const uint32_t flags = (uint32_t)(U(rng) % 4);
    const int64_t hw = (int64_t)H * W;

    std::vector<int32_t> qmask(hw);
    // Gram work grows as bins x (N / tile)^2: many bins go with short series.
    const int pattern = (N > 4097) ? 4 : U(rng) % 6;
    const int nb = 1 + U(rng) % (pattern == 2 ? 3000 : N > 20000 ? 3 : N > 4097 ? 12 : 40);
    for (int64_t i = 0; i < hw; ++i) {
        switch (pattern) {
            case 0: qmask[i] = 1 + U(rng) % nb; break;                 // dense bins
            case 1: qmask[i] = (U(rng) % 5 == 0) ? 1 + U(rng) % nb : 0; break;  // sparse
            case 2: qmask[i] = 1 + U(rng) % nb; break;                 // many bins
            case 3: qmask[i] = (i == hw / 2) ? nb : -(U(rng) % 3); break;  // one pixel
            case 4: qmask[i] = 1 + (int32_t)((i % W) * nb / W); break;     // column rings
            default: qmask[i] = (U(rng) % 2) ? 3 * (1 + U(rng) % nb) : 0; break;  // gaps
        }
    }
    const bool use_mask = U(rng) % 8 != 0;
    int32_t n_bins = 1;
    std::vector<int64_t> npix;
    if (use_mask) {
        n_bins = 0;
        for (int32_t q : qmask) n_bins = q > n_bins ? q : n_bins;
        npix.assign(n_bins, 0);
        for (int32_t q : qmask)
            if (q > 0) ++npix[q - 1];
    } else {
        npix.assign(1, hw);
    }
    std::vector<double> scale;
    const bool use_scale = U(rng) % 4 == 0;
    if (use_scale)
        for (int b = 0; b < n_bins; ++b) scale.push_back(1.0 + b);

    dxg2_plan_desc d{H, W, N, dtype, use_mask ? qmask.data() : nullptr,
                     use_scale ? scale.data() : nullptr, (int32_t)scale.size(), flags};
    int64_t s[12] = {0};
    int rc = dxg2_plan_dry_run(&d, s, 12);
    if (rc != DXG2_OK) return fail(c, "status", rc, DXG2_OK);

    int64_t padded = 0;
    for (int64_t n : npix) padded += (n + kBlock - 1) / kBlock * kBlock;
    const int64_t planes = dtype == DXG2_U16 ? 2 : 1;
    const int64_t nrows = (N + kRowAlign - 1) / kRowAlign * kRowAlign;
    if (s[0] != n_bins) return fail(c, "n_bins", s[0], n_bins);
    if (s[1] != padded) return fail(c, "padded pixels", s[1], padded);
    if (s[2] != planes * padded * nrows) return fail(c, "slab bytes", s[2], planes * padded * nrows);
    for (int i = 3; i < 12; ++i)
        if (s[i] < 0) return fail(c, "work list size", s[i], 0);
    if (padded > 0 && s[3] == 0) return fail(c, "gather items", s[3], 1);
    if (padded > 0 && s[7] == 0) return fail(c, "Gram items", s[7], 1);
    // the band gather, when it has bands, covers every k-block once: its own list plus the
    // fallback's
    if (s[4] > 0 && s[5] + s[6] != padded / kBlock)
        return fail(c, "band + fallback k-blocks", s[5] + s[6], padded / kBlock);

    // the same description with a wrong number of scale factors is refused
    if (n_bins > 1) {
        std::vector<double> short_scale(n_bins - 1, 1.0);
        dxg2_plan_desc e = d;
        e.scale_factors = short_scale.data();
        e.n_scale_factors = n_bins - 1;
        rc = dxg2_plan_dry_run(&e, s, 12);
        if (rc != DXG2_EINVAL) return fail(c, "short scale status", rc, DXG2_EINVAL);
    }
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const int cases = argc > 1 ? std::atoi(argv[1]) : 200;
    std::mt19937_64 rng(12345);
    int bad = 0;
    for (int c = 0; c < cases; ++c) bad += check_case(c, rng);

    // malformed descriptions
    int32_t one = 1;
    dxg2_plan_desc z{0, 4, 10, DXG2_U8, nullptr, nullptr, 0, 0};
    bad += dxg2_plan_dry_run(&z, nullptr, 0) != DXG2_EINVAL;
    z = dxg2_plan_desc{4, 4, 0, DXG2_U8, nullptr, nullptr, 0, 0};
    bad += dxg2_plan_dry_run(&z, nullptr, 0) != DXG2_EINVAL;
    z = dxg2_plan_desc{4, 4, 10, 7, nullptr, nullptr, 0, 0};
    bad += dxg2_plan_dry_run(&z, nullptr, 0) != DXG2_EINVAL;
    z = dxg2_plan_desc{1, 1, 10, DXG2_U8, &one, nullptr, 0, 0};
    bad += dxg2_plan_dry_run(&z, nullptr, 0) != DXG2_OK;
    bad += dxg2_plan_dry_run(nullptr, nullptr, 0) != DXG2_EINVAL;
    z = dxg2_plan_desc{65536, 65536, 10, DXG2_U8, nullptr, nullptr, 0, 0};
    bad += dxg2_plan_dry_run(&z, nullptr, 0) != DXG2_EINVAL;

    std::printf("%d cases, %d failed\n", cases, bad);
    return bad ? 1 : 0;
}
