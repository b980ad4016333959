// gta_slots.h -- slot groups of a few lanes for the superposition-refinement kernels.
//
// A refinement kernel works on many independent slots per pair (a seed
// diagonal, or a fragment length and position); each slot iterates
// covariance pass -> 3x3 Kabsch solve -> TM-score pass.  The solve is a long
// serial chain computed once per group, so one wave runs 64/G slots at once,
// one per group of G lanes (a DPP row or part of one): the passes stride
// positions over the group's lanes, the group reduces with log2(G) DPP steps
// (every lane ends with the same, bitwise identical sum), and the groups'
// solves run side by side in the same instructions.
#pragma once

#include "gta_device.h"

namespace gta {

// 8 lanes by default: measured on the 16 x 50,000 workload, DPRefine passes
// 0.897 s with 16 lanes, 0.840 s with 8, 1.108 s with 4 (gapless refinement
// 0.195 / 0.185 / 0.246 s)
#ifndef GTA_SLOT_LANES
#define GTA_SLOT_LANES 8
#endif
constexpr int GL = GTA_SLOT_LANES;      // lanes per slot group: a DPP row (16) or half row (8)
constexpr int NGW = GTA_WAVE / GL;      // slot groups per wave
static_assert(GL == 8 || GL == 16, "the refinement kernels' groups are a DPP row or half row");

// Every group-level helper takes the group width G (1, 2, 4, 8 or 16 lanes) as
// a template parameter, GL by default; a kernel may use its own width.
template <int G = GL> __device__ __forceinline__ int glane() { return lane_id() & (G - 1); }

// DPP row permutations used by the group reductions
enum { DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_MIRROR = 0x140, DPP_HMIRROR = 0x141 };

template <int CTRL> __device__ __forceinline__ int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL> __device__ __forceinline__ float dpp_f(float v) {
    return __builtin_bit_cast(float, dpp_i<CTRL>(__builtin_bit_cast(int, v)));
}
template <int CTRL> __device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_i<CTRL>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_i<CTRL>((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | (unsigned long long)lo);
}

// Sum over the G lanes of the group.  Partners at each step hold disjoint
// partial sums and add them commutatively, so all lanes end bitwise equal.
template <int G = GL> __device__ __forceinline__ float gsum(float v) {
    if (G >= 2) v = v + dpp_f<DPP_XOR1>(v);
    if (G >= 4) v = v + dpp_f<DPP_XOR2>(v);
    if (G >= 8) v = v + dpp_f<DPP_HMIRROR>(v);
    if (G >= 16) v = v + dpp_f<DPP_MIRROR>(v);
    return v;
}
template <int G = GL> __device__ __forceinline__ double gsum(double v) {
    if (G >= 2) v = v + dpp_d<DPP_XOR1>(v);
    if (G >= 4) v = v + dpp_d<DPP_XOR2>(v);
    if (G >= 8) v = v + dpp_d<DPP_HMIRROR>(v);
    if (G >= 16) v = v + dpp_d<DPP_MIRROR>(v);
    return v;
}
template <int G = GL> __device__ __forceinline__ int gsum(int v) {
    if (G >= 2) v = v + dpp_i<DPP_XOR1>(v);
    if (G >= 4) v = v + dpp_i<DPP_XOR2>(v);
    if (G >= 8) v = v + dpp_i<DPP_HMIRROR>(v);
    if (G >= 16) v = v + dpp_i<DPP_MIRROR>(v);
    return v;
}

// three smallest of the row's values: multiset (M = Min3) or distinct (M = Min3d)
template <int CTRL, class M> __device__ __forceinline__ void gmerge3_step(M& m) {
    const float x = dpp_f<CTRL>(m.a), y = dpp_f<CTRL>(m.b), z = dpp_f<CTRL>(m.c);
    m.put(x);
    m.put(y);
    m.put(z);
}
template <int G = GL, class M> __device__ __forceinline__ float gthird(M& m) {
    if (G >= 2) gmerge3_step<DPP_XOR1>(m);
    if (G >= 4) gmerge3_step<DPP_XOR2>(m);
    if (G >= 8) gmerge3_step<DPP_HMIRROR>(m);
    if (G >= 16) gmerge3_step<DPP_MIRROR>(m);
    return m.c;
}

// covariance accumulators reduced over the row
struct GAccF32 {
    float s[15];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int i = 0; i < 15; i++) s[i] = 0.f;
    }
    __device__ __forceinline__ void add(float qx, float qy, float qz, float rx, float ry, float rz) {
        s[0] += qx * rx; s[1] += qx * ry; s[2] += qx * rz;
        s[3] += qy * rx; s[4] += qy * ry; s[5] += qy * rz;
        s[6] += qz * rx; s[7] += qz * ry; s[8] += qz * rz;
        s[9] += qx; s[10] += qy; s[11] += qz;
        s[12] += rx; s[13] += ry; s[14] += rz;
    }
    template <int G = GL> __device__ __forceinline__ void reduce(float* out) {
#pragma unroll
        for (int i = 0; i < 15; i++) out[i] = gsum<G>(s[i]);
    }
};
// FP64 per lane: products of FP32 coordinates are exact in FP64 (24+24 < 53
// significand bits) and the sums carry FP64 rounding, so the covariance and
// the sums of squares that the RMSD's E0 identity differences are accurate to
// ~1e-16 relative.  (Kahan-compensated FP32 sums still round every product to
// FP32 first: on near-exact copies that leaves RMSD errors of several 1e-3 A.)
struct GAccF64 {
    double s[15];
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int i = 0; i < 15; i++) s[i] = 0.0;
    }
    __device__ __forceinline__ void add(float qx, float qy, float qz, float rx, float ry, float rz) {
        const double x = qx, y = qy, z = qz, a = rx, b = ry, c = rz;
        s[0] += x * a; s[1] += x * b; s[2] += x * c;
        s[3] += y * a; s[4] += y * b; s[5] += y * c;
        s[6] += z * a; s[7] += z * b; s[8] += z * c;
        s[9] += x; s[10] += y; s[11] += z;
        s[12] += a; s[13] += b; s[14] += c;
    }
    template <int G = GL> __device__ __forceinline__ void reduce(double* out) {
#pragma unroll
        for (int i = 0; i < 15; i++) out[i] = gsum<G>(s[i]);
    }
};
template <int P> struct GScore;
template <> struct GScore<0> {
    float s;
    __device__ __forceinline__ void clear() { s = 0.f; }
    __device__ __forceinline__ void add(float v) { s += v; }
    template <int G = GL> __device__ __forceinline__ float reduce() { return gsum<G>(s); }
};
template <> struct GScore<1> {
    float s, c;
    __device__ __forceinline__ void clear() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float v) {
        const float y = __fsub_rn(v, c), t = __fadd_rn(s, y);
        c = __fsub_rn(__fsub_rn(t, s), y);
        s = t;
    }
    template <int G = GL> __device__ __forceinline__ float reduce() { return (float)gsum<G>((double)s - (double)c); }
};

template <int P> struct GPrec;
template <> struct GPrec<0> { typedef float T; typedef GAccF32 Acc; };
template <> struct GPrec<1> { typedef double T; typedef GAccF64 Acc; };

// Groups of this wave that need a slot take consecutive indices from the
// workgroup's counter in LDS.  Called by the whole wave (convergent).
template <int G = GL> __device__ __forceinline__ int fetch_slot(bool need, int* ctr) {
    const unsigned long long lead = __ballot(need && glane<G>() == 0);
    const int cnt = __popcll(lead);
    int base = 0;
    if (cnt) {
        int b = 0;
        if (lane_id() == 0) b = atomicAdd(ctr, cnt);
        base = __shfl(b, 0, GTA_WAVE);
    }
    const int g0 = lane_id() & ~(G - 1);
    const int rank = __popcll(lead & ((1ull << g0) - 1ull));
    return need ? base + rank : -1;
}

}  // namespace gta
