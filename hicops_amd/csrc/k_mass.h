// k_mass.h -- residue masses and the peptide precursor mass on the device,
// shared by the index build (k_index.hip) and the peptide database stage
// (k_lbe.hip).
#pragma once
#include "hcx_internal.h"

namespace hcx {
namespace {

// utils.cpp:41-98 residue masses (float) + static C mod
__constant__ float c_aa[26] = {
    71.03712f, -20000.0f, 103.00919f, 115.030f, 129.0426f, 147.068f, 57.02146f,
    137.060f, 113.084f, -20000.0f, 128.094f, 113.084f, 131.0405f, 114.043f, -20000.0f,
    97.0527f, 128.05858f, 156.1012f, 87.032f, 101.0476f, -20000.0f, 99.06841f,
    186.0793f, -20000.0f, 163.0633f, -20000.0f};

__device__ inline float aa_f(char c) {
    float m = c_aa[c - 'A'];
    float s = (c == 'C') ? 57.021464f : 0.0f;
    return m + s;
}
// GETAA(x, z) = float(AAMass + StatMods) + PROTON * z, in double (utils.cpp:101)
__device__ inline double getaa(char c, uint32_t z) {
    return (double)aa_f(c) + 1.00727647 * (double)z;
}

__device__ inline float pep_mass_dev(const char *seq, uint32_t L) {
    float m = 18.015f;
    for (uint32_t l = 0; l < L; l++) {
        char c = seq[l];
        m += c_aa[c - 'A'];
        m += (c == 'C') ? 57.021464f : 0.0f;
    }
    return m;
}

}  // namespace
}  // namespace hcx
