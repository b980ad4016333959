// gta_device.h -- device-side data model and math of the alignment stage.
//
// Arithmetic that decides DP alignments (transform, squared distance, pair
// score) is written operation by operation in the order the reference CPU
// path evaluates it (transformbase.h:40-109, scoringbase.h:61-65), and the
// translation unit is compiled with FP contraction off, so that for identical
// superpositions the DP matrices -- and hence the alignments -- are bit-exact.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GTA_WAVE 64

namespace gta {

// ---- constants (reference file:line) ----------------------------------------
constexpr float CP_LARGEDST = 999999.9f;          // stagecnsts.cuh:17
constexpr float CP_LARGEDST_cmp = 399999.9f;      // stagecnsts.cuh:19
constexpr float RM_CONVEPSILON = 1.e-5f;          // stagecnsts.cuh:42
constexpr float SCORE_CONVEPSILON = 1.e-6f;       // scoringbase.h:36
constexpr float TFMEPSILON = 1.e-8f;              // transformbase.h:16
constexpr float TFMTOL = 0.01f;                   // transformbase.h:17
constexpr int NMAXSUBFRAGS = 6;                   // mpprocconfbase.h:28
constexpr int SFRAGSTEP = 40;                     // mpprocconfbase.h:30
constexpr int NMAXCONVIT = 20;                    // mpprocconfbase.h:33
constexpr float DEFCOORD_QRY_cmp = 99990.0f;      // mpdpbase.h:26
constexpr float DEFCOORD_RFN_cmp = -99990.0f;     // mpdpbase.h:29
constexpr float EQUIV_D2 = 25.0f;                 // gproc.h:10, squared

enum { CONV_SCOREDP = 2, CONV_LOWTMSC = 8 };      // gproc.h:39-49
enum { BT_STOP = 0, BT_LEFT = 1, BT_UP = 2, BT_DIAG = 3 };  // gproc.h:180-185

// covariance record: 9 cross terms, query sum, reference sum, count
enum { CCM00 = 0, CVq0 = 9, CVr0 = 12, NALN = 15, NCC = 16 };

// waves per workgroup of the refinements (one pair, or one part of a pair,
// per workgroup): two waves keep the share of a workgroup's time spent
// waiting for its last slots small and leave LDS to stage alignments of up to
// 682 positions
constexpr int REFINE_WAVES = 2;
// workgroups of the refinements resident per CU (4 waves per SIMD)
constexpr int REFINE_WGS_PER_CU = 16 / REFINE_WAVES;
// Alignments staged in LDS by the FP32 refinements (24 B per position) in
// workgroups of W waves, 16 waves per CU: the workgroups' fixed LDS (just
// under 2 KB per wave) and staged alignments fit the CU's 160 KB (W = 2: 682
// positions; W = 1: 341).  Longer alignments are read from global memory.
constexpr int rlds_cap32(int W) { return (160 * 1024 * W / 16 - 2048 * W) / 24; }
// waves per workgroup of the seed search (every pair's slots are many and
// even, so a larger workgroup costs no idle lanes)
constexpr int SEARCH_WAVES = 4;

// ---- chunk and pair-state views passed to every kernel -----------------------
struct DevChunk {
    int nstrs, nposs;
    const int* len;
    const unsigned* dst;
    const int* na;          // per structure: 1 = nucleic acid
    const float* x;
    const float* y;
    const float* z;
    const char* rsd;
    char* ss;
    int* napos;             // per position: nucleic-acid pairing partner (scratch of k_secstr)
};

struct Geom {
    int n1, stepinit, nseeds;   // seed diagonals n1 + k*stepinit, k < nseeds (MpStage1.cpp:286-296)
    int nlocsteps;              // refinement slots (MpStage1.cpp:512-516)
    int maxnfragfcts;           // final-stage positions (MpStageFin.cpp:308-309)
};

// A device work queue: items[0 .. *count) taken in order by the waves or
// workgroups of a persistent launch through the counter *head (zeroed
// before the launch).  items == nullptr: no queue (one block per pair).
struct WorkQ {
    const int* items;
    const int* count;
    int* head;
};

struct DevState {
    int nq, nr;
    int* flags;
    float* grand;
    float* best0;
    float* prevbest0;
    int* nalnposs;
    int* qrypos;
    int* rfnpos;
    int* sfragndx;
    int* sfragpos;
    float* tfm_ibest;       // [pair][12]
    float* tfm_grand;       // [pair][12]
    float* out;             // [pair][15]
    // matched pairs, reverse alignment order, region of pair (q,r) starts at
    // q*nrposs + dst[r]; planes: qx qy qz rx ry rz d2 qpos rpos
    float* aln;
    size_t alnplane;
    unsigned* btck;         // packed backtrace codes
    const unsigned long long* btckoff;  // per-pair offsets (in words)
    float* rowbuf;          // strip boundary rows for long references, laid out like aln planes
    unsigned long long* cells;   // DP cells: [0] useful (Lq x Lr per fill), [1] launched lane-steps
    unsigned long long* algbytes;  // algorithmic bytes per kernel class (KC_*)
    unsigned long long* algflops;  // algorithmic FP32 operations per kernel class
};

// kernel classes for the measurement counters
enum { KC_SECSTR = 0, KC_INIT, KC_FINDFRAG, KC_REFINE_GAPLESS, KC_DP, KC_REFINE_DP, KC_REFINE_FIN,
       KC_PROD_ALN, KC_MISC, KC_COUNT };

// algorithmic bytes per aligned pair and pass, and per slot iteration
// (covariance pass 24 B, TM-score pass and thresholded covariance pass 28 B,
// RMSD pass 24 B; 64 B covariance record + 48 B transform per iteration)
constexpr unsigned long long AB_CC = 24, AB_TM = 28, AB_EXT = 28, AB_RMSD = 24, AB_SLOT = 112;
// algorithmic FP32 operations per aligned pair and pass, per slot iteration and
// per DP cell: covariance update 24 (9 products + 15 sums), thresholded
// covariance 25 (one compare more), TM-score pass 33 (transform 21, distance
// 9, score 3), Kabsch solve with its transform ~200, DP cell 12 (3
// subtractions, 3 multiply-adds, an add, a divide, an add, 2 max, a sign)
constexpr unsigned long long AF_CC = 24, AF_EXT = 25, AF_TM = 33, AF_SLOT = 200, AF_CELL = 12;

__device__ __forceinline__ float sqrd(float x) { return x * x; }

// ---- fragment geometry (fragment.cuh:19-109) --------------------------------
__device__ __forceinline__ int GetFragLength(int qlen, int rlen, int qpos, int rpos, int ndx) {
    int aln = min(qlen - qpos, rlen - rpos);
    if (aln < 4 && ndx == 0) return aln;
    aln >>= ndx;
    if (aln == 2 && 0 < ndx && ((aln >> (ndx - 1)) & 1)) aln = 4;
    if (aln < 3) return 0;
    if (aln < 4) aln = 4;
    if (ndx + 1 == NMAXSUBFRAGS) aln = 4;
    return aln;
}
__device__ __forceinline__ void UpdateLengths(int& qlen, int& rlen, int& qpos, int& rpos, int fl) {
    int qe = qpos + fl, re = rpos + fl;
    int delta = max(qe - qlen, re - rlen);
    if (0 < delta) { qpos -= delta; rpos -= delta; }
    else { qlen = qe; rlen = re; }
}

// ---- transform / distance / score, reference evaluation order --------------
__device__ __forceinline__ void tfm_point(const float* t, float& x0, float& x1, float& x2) {
    float x = x0, y = x1, z = x2;
    x0 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(t[0], x), __fmul_rn(t[1], y)), __fmul_rn(t[2], z)), t[9]);
    x1 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(t[3], x), __fmul_rn(t[4], y)), __fmul_rn(t[5], z)), t[10]);
    x2 = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(t[6], x), __fmul_rn(t[7], y)), __fmul_rn(t[8], z)), t[11]);
}
__device__ __forceinline__ float dist2(float a0, float a1, float a2, float b0, float b1, float b2) {
    float d0 = __fsub_rn(a0, b0), d1 = __fsub_rn(a1, b1), d2 = __fsub_rn(a2, b2);
    return __fadd_rn(__fadd_rn(__fmul_rn(d0, d0), __fmul_rn(d1, d1)), __fmul_rn(d2, d2));
}
__device__ __forceinline__ float tfm_dist2(const float* t, float qx, float qy, float qz, float rx,
                                           float ry, float rz) {
    tfm_point(t, qx, qy, qz);
    return dist2(qx, qy, qz, rx, ry, rz);
}
// Correctly rounded a / b for positive normal a and b whose quotient is
// normal and whose magnitudes are far from the exponent limits (here
// a = d0^2 in [0.9, 2.4e3], b = a + d^2 < 2^100).  It is the Newton-Raphson
// sequence with FMA of IEEE division on this GPU (reciprocal refined once,
// two residual corrections of the quotient) without the range scaling and
// special-case fix-up steps, which leave such operands unchanged: 8 VALU
// instructions instead of 11, bit-identical results.
__device__ __forceinline__ float div_rn_pos(float a, float b) {
    float y = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = __fmul_rn(a, y);
    float r = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(r, y, q);
    r = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(r, y, q);
}
// d0^2 / (d0^2 + d^2), correctly rounded as the reference's IEEE division
__device__ __forceinline__ float pair_score(float d02, float d2) {
    return div_rn_pos(d02, __fadd_rn(d02, d2));
}

// ---- wave64 collectives -------------------------------------------------------
__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

__device__ __forceinline__ int wave_sum_i(int v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, GTA_WAVE);
    return v;
}

// three smallest values of a multiset, kept sorted per lane
// (a <= b <= c: the lowest three of {a, b, c, d} are min(a, d), the median of
// {a, b, d} and the median of {b, c, d}, the same triple as sorted insertion;
// three v_med3_f32 without branches, and, unlike fminf/fmaxf, without the
// operand canonicalisations the compiler adds for IEEE mode.  Distances are
// never NaN, so the medians equal the min/max network's values)
struct Min3 {
    float a, b, c;
    __device__ void init() { a = b = c = CP_LARGEDST; }
    __device__ __forceinline__ void put(float d) {
        const float na = __builtin_amdgcn_fmed3f(a, d, -__builtin_inff());
        const float nb = __builtin_amdgcn_fmed3f(a, b, d);
        c = __builtin_amdgcn_fmed3f(b, c, d);
        b = nb;
        a = na;
    }
};
// three smallest distinct values (a value already held changes nothing: it
// enters as c, which leaves the triple as it is)
struct Min3d {
    float a, b, c;
    __device__ void init() { a = b = c = CP_LARGEDST; }
    __device__ __forceinline__ void put(float d) {
        if (d == a || d == b || d == c) d = c;
        const float na = __builtin_amdgcn_fmed3f(a, d, -__builtin_inff());
        const float nb = __builtin_amdgcn_fmed3f(a, b, d);
        c = __builtin_amdgcn_fmed3f(b, c, d);
        b = nb;
        a = na;
    }
};

}  // namespace gta
