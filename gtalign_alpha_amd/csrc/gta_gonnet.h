// gta_gonnet.h -- pairwise residue scores of the sequence-and-structure search.
//
// The --add-search-by-ss alignment scores residue pairs with the Gonnet,
// Cohen & Benner (1992, Science 256:1443) mutation matrix, in bits: the
// reference derives its table (libutil/alpha.cpp:167-206) as
// log2(10^(G/10)), with the frequency ratio 10^(G/10) kept to six decimals,
// G the published matrix entry (tenths of a dB, one decimal).  The 20 standard
// amino acids carry G below; every pair involving another letter of the
// alphabet (B, J, O, U, X, Z) has ratio 1 and score 0, and non-letters score 0
// at lookup.  Host code only: the table is built once with the C library's
// logf, as the reference does, and uploaded to the device.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace gta {

constexpr int GONNET_NA = 26;   // table over the English alphabet

// lower triangle of the Gonnet matrix for ACDEFGHIKLMNPQRSTVWY
inline const double* gonnet_lower() {
    static const double g[] = {
        2.4,
        0.5, 11.5,
        -0.3, -3.2, 4.7,
        0.0, -3.0, 2.7, 3.6,
        -2.3, -0.8, -4.5, -3.9, 7.0,
        0.5, -2.0, 0.1, -0.8, -5.2, 6.6,
        -0.8, -1.3, 0.4, 0.4, -0.1, -1.4, 6.0,
        -0.8, -1.1, -3.8, -2.7, 1.0, -4.5, -2.2, 4.0,
        -0.4, -2.8, 0.5, 1.2, -3.3, -1.1, 0.6, -2.1, 3.2,
        -1.2, -1.5, -4.0, -2.8, 2.0, -4.4, -1.9, 2.8, -2.1, 4.0,
        -0.7, -0.9, -3.0, -2.0, 1.6, -3.5, -1.3, 2.5, -1.4, 2.8, 4.3,
        -0.3, -1.8, 2.2, 0.9, -3.1, 0.4, 1.2, -2.8, 0.8, -3.0, -2.2, 3.8,
        0.3, -3.1, -0.7, -0.5, -3.8, -1.6, -1.1, -2.6, -0.6, -2.3, -2.4, -0.9, 7.6,
        -0.2, -2.4, 0.9, 1.7, -2.6, -1.0, 1.2, -1.9, 1.5, -1.6, -1.0, 0.7, -0.2, 2.7,
        -0.6, -2.2, -0.3, 0.4, -3.2, -1.0, 0.6, -2.4, 2.7, -2.2, -1.7, 0.3, -0.9, 1.5, 4.7,
        1.1, 0.1, 0.5, 0.2, -2.8, 0.4, -0.2, -1.8, 0.1, -2.1, -1.4, 0.9, 0.4, 0.2, -0.2, 2.2,
        0.6, -0.5, 0.0, -0.1, -2.2, -1.1, -0.3, -0.6, 0.1, -1.3, -0.6, 0.5, 0.1, 0.0, -0.2, 1.5, 2.5,
        0.1, 0.0, -2.9, -1.9, 0.1, -3.3, -2.0, 3.1, -1.7, 1.8, 1.6, -2.2, -1.8, -1.5, -2.0, -1.0,
        0.0, 3.4,
        -3.6, -1.0, -5.2, -4.3, 3.6, -4.0, -0.8, -1.8, -3.5, -0.7, -1.0, -3.6, -5.0, -2.7, -1.6,
        -3.3, -3.5, -2.6, 14.2,
        -2.2, -0.5, -2.8, -2.7, 5.1, -4.0, 2.2, -0.7, -2.1, 0.0, -0.2, -1.4, -3.1, -1.7, -1.8,
        -1.9, -1.9, -1.1, 4.1, 7.8,
    };
    return g;
}

// bits[(a-'A')*26 + (b-'A')]: log2 of the frequency ratio, as single precision
inline void gonnet_bits(float* bits) {
    static const char aa[] = "ACDEFGHIKLMNPQRSTVWY";
    float ratio[GONNET_NA * GONNET_NA];
    for (int k = 0; k < GONNET_NA * GONNET_NA; k++) ratio[k] = 1.0f;
    const double* g = gonnet_lower();
    for (int i = 0, k = 0; i < 20; i++)
        for (int j = 0; j <= i; j++, k++) {
            // the ratio as a six-decimal literal, parsed to single precision
            char buf[32];
            snprintf(buf, sizeof buf, "%.6f", pow(10.0, g[k] / 10.0));
            const float r = strtof(buf, nullptr);
            const int a = aa[i] - 'A', b = aa[j] - 'A';
            ratio[a * GONNET_NA + b] = r;
            ratio[b * GONNET_NA + a] = r;
        }
    const float ln2_recip = 1.44269504088896338700f;
    for (int k = 0; k < GONNET_NA * GONNET_NA; k++) bits[k] = logf(ratio[k]) * ln2_recip;
}

}  // namespace gta
